# Builds the reference's own stack I/O (src-1125/mrc.cpp, mrcz.cpp, mrct.cpp,
# compiled where they lie under /root/reference) with a C shim into
# oracle/_ref/libref_io.so.  The rest of the reference path (DFAlign.cu,
# cufunc.cu, func.cpp, mfunc.h) needs CUDA, cuFFT, FFTW and SP++, none of
# which exist here, so it is not built.  libtiff (conda) is linked by path;
# tests preload it.
REF ?= /root/reference/src-1125
TIFF_PREFIX ?= /opt/conda
_ref/libref_io.so: ref_io_shim.cpp $(REF)/mrc.cpp $(REF)/mrcz.cpp $(REF)/mrct.cpp
	mkdir -p _ref
	g++ -O2 -fPIC -shared -w -fpermissive -I$(REF) -I$(TIFF_PREFIX)/include \
	    ref_io_shim.cpp $(REF)/mrc.cpp $(REF)/mrcz.cpp $(REF)/mrct.cpp \
	    $(TIFF_PREFIX)/lib/libtiff.so.5 -Wl,-rpath-link,$(TIFF_PREFIX)/lib -lz -o $@
