# Builds the reference AOCL-FFTZ CPU library from its own sources, where they
# lie, into oracle/_ref/ (never into the reference tree), and the reference's
# own benchmark program (test/bench, test/utils/dims_vecs_helper.c; target
# aocl_fftz_bench of CMakeLists.txt:589-600) linked against that library.  The object list and
# per-directory ISA flags follow the reference's CMakeLists.txt:70-175 (source
# globs) and :200-242 (AVX flags), with ENABLE_MULTI_THREADING=ON (:247-265).
# Nothing from the reference is copied into this repository; the resulting
# .so is test infrastructure only (cross-check and CPU baseline).
#
#   make -f oracle/ref.mk -j8            (from the repository root)
#
REF      ?= /root/reference
OUT      := oracle/_ref
OBJDIR   := $(OUT)/obj
CC       := gcc

SRCS := $(shell find $(REF)/api $(REF)/core $(REF)/selector $(REF)/utils -name '*.c' | sort)
OBJS := $(patsubst $(REF)/%.c,$(OBJDIR)/%.o,$(SRCS))

DEFS := -DNDEBUG -DENABLE_AVX512 -DENABLE_AVX256 -DENABLE_AVX128 -DENABLE_FMA \
        -DMULTI_THREADING -DREAL_FFT_EXECUTION_ORDER=2 \
        -DAOCL_BUILD_VERSION='"Build oracle-ref"'
BASE := -O3 -fPIC -fomit-frame-pointer -fstrict-aliasing -fopenmp -w \
        -I$(REF) -I$(REF)/api $(DEFS)

isa_flags = $(if $(findstring /avx512/,$1),-mavx512f -mavx512dq -ffp-contract=fast,\
            $(if $(findstring /avx256/,$1),-mavx -mfma -ffp-contract=fast,\
            $(if $(findstring /avx128/,$1),-mavx -mfma -ffp-contract=fast,\
            $(if $(findstring /kernels/,$1),-fno-version-loops-for-strides,))))

BENCH_SRCS := $(sort $(wildcard $(REF)/test/bench/*.c $(REF)/test/bench/utils/*.c)) $(REF)/test/utils/dims_vecs_helper.c
BENCH_OBJS := $(patsubst $(REF)/%.c,$(OBJDIR)/%.o,$(BENCH_SRCS))
BENCH_INC  := -I$(REF)/test/bench -I$(REF)/test/bench/utils -I$(REF)/test/utils

all: $(OUT)/libaocl_fftz_ref.so $(OUT)/aocl_fftz_bench

$(OBJDIR)/%.o: $(REF)/%.c
	@mkdir -p $(dir $@)
	@$(CC) $(BASE) $(call isa_flags,$<) -c $< -o $@

# -Bsymbolic keeps the library's calls to its own aoclfftz_* entry points
# inside itself when a process also loads the GPU library (same names).
$(OUT)/libaocl_fftz_ref.so: $(OBJS)
	$(CC) -shared -fopenmp -Wl,-Bsymbolic -o $@ $(OBJS) -lm

# the benchmark: timing build (ACCURACY_WITH_DFT off), finds the library
# beside itself
$(OBJDIR)/test/%.o: $(REF)/test/%.c
	@mkdir -p $(dir $@)
	@$(CC) $(BASE) $(BENCH_INC) -c $< -o $@

$(OUT)/aocl_fftz_bench: $(BENCH_OBJS) $(OUT)/libaocl_fftz_ref.so
	$(CC) -fopenmp -o $@ $(BENCH_OBJS) -L$(OUT) -laocl_fftz_ref -Wl,-rpath,'$$ORIGIN' -lm -ldl

clean:
	rm -rf $(OUT)

.PHONY: all clean
