# Compiles the reference's kirchmig2d_cig path (and the table/shot/compare
# helpers, the layer stitch/extract programs and the segmented butterfly used
# to write golden fixtures) directly from its sources under
# REF, with the reference's release flags (-O3 -DNDEBUG + OpenMP).  Nothing
# is copied, generated or stubbed: the vendored fftw3.h header and sqlite3.c
# source are the reference's own.  Outputs go to _ref/ only.
REF ?= /root/reference
OUT := _ref
CXX ?= g++
CXXFLAGS := -O3 -DNDEBUG -fopenmp -DSE_USE_OMP -w -I$(REF) -I$(REF)/SEBASIC/include \
  -I$(REF)/SEFILESYSTEM/include -I$(REF)/SERECKIRCH/include -I$(REF)/fftw/api
SB := $(shell sed -n '/^set(SEBASIC_SOURCES/,/^)/p' $(REF)/SEBASIC/CMakeLists.txt | grep -o 'src/[a-z_0-9]*\.cpp')
SF := $(shell sed -n '/^set(SEFS_SOURCES/,/^)/p' $(REF)/SEFILESYSTEM/CMakeLists.txt | grep -o 'src/[a-z_0-9]*\.cpp')
SR := kirmig.cpp tinterpsr.cpp eikods.cpp neighbors.cpp pqueue.cpp decart.cpp efmm.cpp mutter.cpp tpow.cpp kirdat.cpp
OB := $(patsubst %.cpp,$(OUT)/obj/SEBASIC/%.o,$(SB)) $(OUT)/obj/sqlite3.o
OF := $(patsubst %.cpp,$(OUT)/obj/SEFILESYSTEM/%.o,$(SF))
OR := $(patsubst %.cpp,$(OUT)/obj/SERECKIRCH/src/%.o,$(SR))
PROGS := kirchmig2d_cig eikods2d shotgen rsf_l2cmp layercom_cig extract_adcig kirchdat2d_auto_cs \
  greensol_auto

all: $(addprefix $(OUT)/,$(PROGS)) $(OUT)/libbf1d_ref.so

# The butterfly of kirchdat2d_freq (SERECKIRCH/src/bf1d.cpp) as a shared
# library: bf1d_ref_driver.cpp includes that source file where it lies and
# adds ctypes entry points.  It uses only the vendored fftw3.h types.
$(OUT)/libbf1d_ref.so: bf1d_ref_driver.cpp $(REF)/SERECKIRCH/src/bf1d.cpp $(REF)/SERECKIRCH/include/bf1d.h
	@mkdir -p $(OUT)
	$(CXX) -O3 -DNDEBUG -w -fPIC -shared -I$(REF)/SERECKIRCH/src -I$(REF)/fftw/api -o $@ $<

$(OUT)/obj/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/sqlite3.o: $(REF)/SEBASIC/src/sqlite3.c
	@mkdir -p $(dir $@)
	gcc -O2 -w -c $< -o $@

$(OUT)/libsb.a: $(OB)
	ar rcs $@ $^
$(OUT)/libsf.a: $(OF)
	ar rcs $@ $^
$(OUT)/librk.a: $(OR)
	ar rcs $@ $^

$(OUT)/%: $(OUT)/obj/src/%.o $(OUT)/librk.a $(OUT)/libsf.a $(OUT)/libsb.a
	$(CXX) -fopenmp -o $@ $< -Wl,--start-group $(OUT)/librk.a $(OUT)/libsf.a $(OUT)/libsb.a -Wl,--end-group -lpthread -ldl

.PRECIOUS: $(OUT)/obj/%.o $(OUT)/obj/src/%.o
.PHONY: all
