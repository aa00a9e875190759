# Builds the reference CPU C ABI (libufsecp) straight from its sources under
# /root/reference with g++, into oracle/_ref/.  Test infrastructure only: the
# result is a cross-check for the oracle restatement and the CPU baseline of
# the benchmark; it never ships.  Flags follow the reference's own Release
# configuration (-O3, x86-64 field assembly with BMI2/ADX, SECP256K1_HAS_ASM),
# with -march=x86-64-v3 (its reproducible-build target) so the same binary
# runs on this host and on the GPU host.
REF      ?= /root/reference
OUT      ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
CXX      ?= g++
CXXFLAGS ?= -O3 -march=x86-64-v3 -std=c++20 -fPIC -w -DNDEBUG -DUFSECP_BUILDING -DSECP256K1_HAS_ASM=1
ASFLAGS  ?= -x assembler-with-cpp -mbmi2 -madx -fPIC
INCS     := -I$(REF)/src/cpu/include -I$(REF)/include -I$(REF)/include/ufsecp -I$(REF)/src/cpu/src
SRCS     := $(filter-out %/ufsecp_gpu_impl.cpp,$(wildcard $(REF)/src/cpu/src/*.cpp))
OBJS     := $(patsubst $(REF)/src/cpu/src/%.cpp,$(OUT)/obj/%.o,$(SRCS)) $(OUT)/obj/field_asm_x64_gas.o

# The reference's own GPU C-ABI audit programs (audit/test_gpu_host_api_negative.cpp,
# audit/test_gpu_abi_gate.cpp), compiled from their sources against this
# repository's include/ufsecp/ufsecp_gpu.h and linked to its libufsecp_gpu.so
# (the reference CPU library above supplies the signing calls they use to make
# inputs): a link-time and run-time check that the drop-in binds as the
# reference's callers expect.
REPO     := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))..)
GPULIB   := $(REPO)/ultrafastsecp256k1_amd/libufsecp_gpu.so
AUDIT    := test_gpu_host_api_negative test_gpu_abi_gate

all: $(OUT)/libufsecp_ref.so $(addprefix $(OUT)/,$(AUDIT))

$(OUT)/test_gpu_%: $(REF)/audit/test_gpu_%.cpp $(OUT)/libufsecp_ref.so $(GPULIB) $(REPO)/include/ufsecp/ufsecp_gpu.h
	$(CXX) -std=c++17 -O1 -w -I$(REPO)/include -I$(REF)/include -I$(REF)/audit $< \
	  -L$(REPO)/ultrafastsecp256k1_amd -lufsecp_gpu -L$(OUT) -lufsecp_ref \
	  -Wl,-rpath,'$$ORIGIN:$$ORIGIN/../../ultrafastsecp256k1_amd' -Wl,-rpath-link,/opt/rocm/lib -o $@

$(OUT)/obj/%.o: $(REF)/src/cpu/src/%.cpp
	@mkdir -p $(OUT)/obj
	$(CXX) $(CXXFLAGS) $(INCS) -c $< -o $@

$(OUT)/obj/field_asm_x64_gas.o: $(REF)/src/cpu/src/field_asm_x64_gas.S
	@mkdir -p $(OUT)/obj
	$(CC) $(ASFLAGS) $(INCS) -c $< -o $@

$(OUT)/libufsecp_ref.so: $(OBJS)
	$(CXX) -shared -o $@ $(OBJS) -lpthread

clean:
	rm -rf $(OUT)

.PHONY: all clean
