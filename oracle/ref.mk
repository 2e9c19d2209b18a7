# Recipe for oracle/_ref: compiles the reference CPU path (the `make CUDA=0`
# configuration: g++ -O3 -fopenmp, .cu files as C++) from its sources where
# they lie, plus the small harness in oracle/ref_harness/.  Outputs go only to
# oracle/_ref/.  Test infrastructure only; never part of the product.
#
#   make -f oracle/ref.mk            (needs /root/reference; skipped without it)
REF ?= /root/reference
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
OBJ := $(OUT)/obj
CXX := g++
# Brotli headers live in /opt/conda/include; -idirafter keeps that directory
# behind the system headers.  Link the system libbrotli by full path.
CXXFLAGS := -g -std=c++17 -O3 -fopenmp -fPIC -w -idirafter /opt/conda/include -I$(REF)
BROTLI := /usr/lib/x86_64-linux-gnu/libbrotlienc.so.1 /usr/lib/x86_64-linux-gnu/libbrotlicommon.so.1

LANGS := bff bff_noheads bff8 bff8_noheads bff_noheads_4bit bff_perm bff_selfmove \
         forth forthcopy forthtrivial forthtrivial_reset subleq rsubleq4
struct_of = $(if $(filter bff%,$(1)),Bff,$(if $(filter forth%,$(1)),Forth,$(if $(filter subleq,$(1)),Subleq,RSubleq4)))

MAIN_OBJS := $(OBJ)/main.o $(OBJ)/common.o $(foreach l,$(LANGS),$(OBJ)/lang_$(l).o)
WRAP_OBJS := $(OBJ)/driver.o $(OBJ)/common.o $(foreach l,$(LANGS),$(OBJ)/wrap_$(l).o)

all: $(OUT)/cubff_ref_main $(OUT)/libcubff_ref.so

$(OBJ):
	mkdir -p $(OBJ)

$(OBJ)/main.o: $(REF)/main.cc | $(OBJ)
	$(CXX) -c $(CXXFLAGS) -xc++ $< -o $@
$(OBJ)/common.o: $(REF)/common.cc | $(OBJ)
	$(CXX) -c $(CXXFLAGS) -xc++ $< -o $@
$(OBJ)/lang_%.o: $(REF)/%.cu | $(OBJ)
	$(CXX) -c $(CXXFLAGS) -xc++ $< -o $@
$(OBJ)/wrap_%.o: $(HERE)ref_harness/lang_wrap.cc | $(OBJ)
	$(CXX) -c $(CXXFLAGS) -xc++ -DREF_LANG_FILE='"$(REF)/$*.cu"' -DREF_LANG_TAG=$* \
	  -DREF_LANG_STRUCT=$(call struct_of,$*) $< -o $@
$(OBJ)/driver.o: $(HERE)ref_harness/driver.cc | $(OBJ)
	$(CXX) -c $(CXXFLAGS) $< -o $@

$(OUT)/cubff_ref_main: $(MAIN_OBJS)
	$(CXX) $^ -fopenmp $(BROTLI) -o $@
$(OUT)/libcubff_ref.so: $(WRAP_OBJS)
	$(CXX) -shared $^ -fopenmp $(BROTLI) -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
