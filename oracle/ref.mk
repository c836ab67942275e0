# Builds the reference OpenZL C library, from its sources where they lie under
# /root/reference, into oracle/_ref/libopenzl_ref.so. Used only as a checker by
# the tests (never shipped, never on the product path). The reference's own
# build system is not used: every src/openzl/**/*.c file plus the x86-64 Huffman
# decode loop is compiled with one gcc line each, and the zstd/lz4 codecs link
# against the zstd 1.4.9 / lz4 1.9.3 libraries that ship in this image
# (/opt/conda). No file of the reference is copied and no stand-in is written.
REF      ?= /root/reference
OUT      := $(dir $(lastword $(MAKEFILE_LIST)))_ref
DEPINC   ?= /opt/conda/include
DEPLIB   ?= /opt/conda/lib
CSRCS    := $(shell find $(REF)/src/openzl -name '*.c' | sort)
SSRCS    := $(REF)/src/openzl/fse/decompress/huf_decompress_amd64.S
OBJS     := $(patsubst $(REF)/%.c,$(OUT)/obj/%.o,$(CSRCS)) \
            $(patsubst $(REF)/%.S,$(OUT)/obj/%.o,$(SSRCS))
CFLAGS   := -std=gnu11 -O2 -fPIC -DNDEBUG -I$(REF)/include -I$(REF)/src -I$(REF) -I$(DEPINC)

all: $(OUT)/libopenzl_ref.so

$(OUT)/obj/%.o: $(REF)/%.c
	@mkdir -p $(dir $@)
	@gcc $(CFLAGS) -c $< -o $@

$(OUT)/obj/%.o: $(REF)/%.S
	@mkdir -p $(dir $@)
	@gcc $(CFLAGS) -c $< -o $@

$(OUT)/libopenzl_ref.so: $(OBJS)
	gcc -shared -o $@ $(OBJS) -L$(DEPLIB) -lzstd -llz4 -Wl,-rpath,$(DEPLIB) -Wl,--no-undefined -lm

.PHONY: all

# Flat ctypes-facing checker around the reference library (test-only).
$(OUT)/libref_harness.so: $(dir $(lastword $(MAKEFILE_LIST)))ref_harness.c $(OUT)/libopenzl_ref.so
	gcc -std=gnu11 -O2 -fPIC -shared -I$(REF)/include -I$(REF)/src -o $@ $< -L$(OUT) -lopenzl_ref -Wl,-rpath,'$$ORIGIN'

all: $(OUT)/libref_harness.so
