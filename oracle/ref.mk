# Builds the reference CPU C ABI (libufsecp) straight from its sources under
# /root/reference with g++, into oracle/_ref/.  Test infrastructure only: the
# result is a cross-check for the oracle restatement and never ships.
REF      ?= /root/reference
OUT      ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
CXX      ?= g++
CXXFLAGS ?= -O2 -std=c++20 -fPIC -w -DNDEBUG -DUFSECP_BUILDING
INCS     := -I$(REF)/src/cpu/include -I$(REF)/include -I$(REF)/include/ufsecp -I$(REF)/src/cpu/src
SRCS     := $(filter-out %/ufsecp_gpu_impl.cpp,$(wildcard $(REF)/src/cpu/src/*.cpp))
OBJS     := $(patsubst $(REF)/src/cpu/src/%.cpp,$(OUT)/obj/%.o,$(SRCS))

all: $(OUT)/libufsecp_ref.so

$(OUT)/obj/%.o: $(REF)/src/cpu/src/%.cpp
	@mkdir -p $(OUT)/obj
	$(CXX) $(CXXFLAGS) $(INCS) -c $< -o $@

$(OUT)/libufsecp_ref.so: $(OBJS)
	$(CXX) -shared -o $@ $(OBJS) -lpthread

clean:
	rm -rf $(OUT)

.PHONY: all clean
