# Builds oracle/_ref/libref.so from the reference's own distiller and scorer
# sources (include/transforms/distiller.hpp, include/transforms/scorer.hpp,
# include/data_types/candidates.hpp: standard C++ only).  The rest of the reference's search path needs CUDA, cuFFT,
# Thrust or libdedisp and is not buildable here.
# Usage: make -C oracle -f ref.mk   (no-op when the reference is absent)
REF ?= /root/reference
CXX ?= g++

_ref/libref.so: ref_driver.cpp
	@if [ -f $(REF)/include/transforms/distiller.hpp ]; then \
	  mkdir -p _ref && \
	  $(CXX) -O2 -std=c++11 -fPIC -shared -ffp-contract=off -w -I$(REF)/include ref_driver.cpp -o $@; \
	else echo "reference sources absent: oracle/_ref not built"; fi

.PHONY: clean
clean:
	rm -rf _ref
