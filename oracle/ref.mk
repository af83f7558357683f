# Builds oracle/_ref/libref.so from the reference's own sources where they lie
# under /root/reference (test infrastructure; see ref_shim.cc).  Only the
# header-only arithmetic and src/sampler/halton.cc: the rest of the path needs
# the reference's CMake-generated export header and is not built.
REF ?= /root/reference
CXX ?= g++
# the reference's default build: FAST_MATH and FAST_TRIG on, x86-64 SSE2
# without FMA contraction
REF_FLAGS = -O2 -std=c++17 -fPIC -ffp-contract=off -fno-fast-math -DFAST_MATH -DFAST_TRIG
_ref/libref.so: ref_shim.cc $(REF)/src/sampler/halton.cc
	mkdir -p _ref
	$(CXX) $(REF_FLAGS) -shared -I$(REF)/include -o $@ ref_shim.cc $(REF)/src/sampler/halton.cc
.PHONY: clean
clean:
	rm -f _ref/libref.so
