# Compiles the reference applications from the reference tree where they lie
# (never copied), with the reference makefile's flags (makefile:10-17), into
# $(OUT)/bin. COMMITS is a literal because the tree has no git metadata.
REF ?= /root/reference
OUT ?= _ref
CFLAGS := -std=c11 -O3 -fopenmp -w
DEFS := -DIP_FORCE_DIRECT_MAPPING -DVERSION=\"1.0.0\" -DIP_MACHINE=\"ref\" -D_GNU_SOURCE \
        -DCOMMITS=\"none\" -DCOMPILATION_FLAGS=\"ref\"

APPS := cc pagerank sssp
VARIANTS := plain spread single_broadcast single_broadcast_spread
BINS := $(OUT)/bin/cc_32 $(OUT)/bin/cc_spread_32 $(OUT)/bin/cc_single_broadcast_32 \
        $(OUT)/bin/cc_single_broadcast_spread_32 \
        $(OUT)/bin/pagerank_32 $(OUT)/bin/pagerank_single_broadcast_32 \
        $(OUT)/bin/sssp_32 $(OUT)/bin/sssp_spread_32 $(OUT)/bin/sssp_single_broadcast_32 \
        $(OUT)/bin/sssp_single_broadcast_spread_32
# the _64 builds (IP_VERTEX_ID_TYPE=uint64_t, makefile:17)
BINS += $(patsubst %_32,%_64,$(BINS))

# benchmarks/pagerank.c with directed=true (its line 91 set to true), built
# from a copy under /tmp that is deleted afterwards: PageRank over the
# in-lists the reference mirrors in ascending source order
# (src/iPregel_postamble.h:429-487)
BINS += $(OUT)/bin/pagerank_directed_single_broadcast_32

all: $(BINS)

$(OUT)/bin/pagerank_directed_single_broadcast_32: | $(OUT)/bin
	tmp=$$(mktemp -d /tmp/ipr_ref_directed_XXXXXX) && \
	sed 's/bool directed = false;/bool directed = true;/' $(REF)/benchmarks/pagerank.c > $$tmp/pagerank.c && \
	grep -q 'bool directed = true;' $$tmp/pagerank.c && \
	gcc $(CFLAGS) $(DEFS) -DIP_APPLICATION=\"pagerank_directed_single_broadcast\" -DIP_USE_SINGLE_BROADCAST \
	    -DIP_VERTEX_ID_TYPE=uint32_t -I$(REF)/src $$tmp/pagerank.c -o $@; \
	rc=$$?; rm -rf $$tmp; exit $$rc

$(OUT)/bin:
	mkdir -p $@

define variant_flags
$(if $(findstring single_broadcast,$(1)),-DIP_USE_SINGLE_BROADCAST) $(if $(findstring spread,$(1)),-DIP_USE_SPREAD)
endef

$(OUT)/bin/%_32: | $(OUT)/bin
	gcc $(CFLAGS) $(DEFS) -DIP_APPLICATION=\"$*\" $(call variant_flags,$*) \
	    -DIP_VERTEX_ID_TYPE=uint32_t -I$(REF)/src \
	    $(REF)/benchmarks/$(firstword $(subst _, ,$*)).c -o $@

$(OUT)/bin/%_64: | $(OUT)/bin
	gcc $(CFLAGS) $(DEFS) -DIP_APPLICATION=\"$*\" $(call variant_flags,$*) \
	    -DIP_VERTEX_ID_TYPE=uint64_t -I$(REF)/src \
	    $(REF)/benchmarks/$(firstword $(subst _, ,$*)).c -o $@
