# Builds the reference renderer (libSLR) from its own sources, where they lie
# under /root/reference, plus this repository's harness around it.
# Output goes only to oracle/_ref/ (git-ignored; it travels with the tree).
#
#   make -f oracle/ref.mk -j8
#
# Flags follow the reference's release build (-O3 -DNDEBUG, gnu++11, no -march,
# so no FMA contraction on x86-64) plus the two portability flags libSLR needs
# on Linux: SLR_Platform_OpenBSD routes SLR_memalign to posix_memalign and
# -include cstring supplies std::memcpy.
REF     ?= /root/reference
SRC      = $(REF)/libSLR
OUT      = oracle/_ref
CXX     ?= g++
CC      ?= gcc
FLAGS    = -O3 -DNDEBUG -fPIC -DSLR_Platform_OpenBSD -w
CXXFLAGS = -std=gnu++11 $(FLAGS) -include cstring -I$(REF)
CFLAGS   = $(FLAGS)

EXCL    := StackAllocator MSpaceAllocator dlmalloc AMCMCPPMRenderer
CPPS    := $(filter-out $(foreach e,$(EXCL),%$(e).cpp),$(shell find $(SRC) -name '*.cpp'))
CS      := $(filter-out $(foreach e,$(EXCL),%$(e).c),$(shell find $(SRC) -name '*.c'))
OBJS    := $(patsubst $(SRC)/%.cpp,$(OUT)/obj/%.o,$(CPPS)) $(patsubst $(SRC)/%.c,$(OUT)/obj/%.o,$(CS))

all: $(OUT)/libslr_ref.so $(OUT)/libslr_refdrv.so $(OUT)/adapter_check

$(OUT)/obj/%.o: $(SRC)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/%.o: $(SRC)/%.c
	@mkdir -p $(dir $@)
	$(CC) $(CFLAGS) -c $< -o $@

$(OUT)/libslr_ref.so: $(OBJS)
	$(CXX) -shared -o $@ $(OBJS) -lpthread

$(OUT)/libslr_refdrv.so: oracle/ref_driver.cpp $(OUT)/libslr_ref.so
	$(CXX) $(CXXFLAGS) -shared -o $@ oracle/ref_driver.cpp -L$(OUT) -lslr_ref -Wl,-rpath,'$$ORIGIN' -lpthread

clean:
	rm -rf $(OUT)

# a libSLR host program using include/slr_hip_renderer.hpp (needs
# slr_amd/_lib/libslr_hip.so, built first by slr_amd/build.py)
# (libslr_refdrv.so supplies the CMF-strata statics libslr_ref.so declares)
$(OUT)/adapter_check: oracle/adapter_check.cpp include/slr_hip_renderer.hpp include/slr_hip.h $(OUT)/libslr_refdrv.so slr_amd/_lib/libslr_hip.so
	$(CXX) $(CXXFLAGS) -Iinclude -o $@ oracle/adapter_check.cpp -L$(OUT) -Wl,--no-as-needed -lslr_refdrv -Wl,--as-needed -lslr_ref -Lslr_amd/_lib -lslr_hip \
		-Wl,-rpath,'$$ORIGIN' -Wl,-rpath,'$$ORIGIN/../../slr_amd/_lib' -lpthread
