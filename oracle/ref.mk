# Builds oracle/_ref/ref_numbering (test infrastructure): the reference's own
# mesh, refinement, FE-space and matrix-structure code compiled with g++ from
# the sources where they lie under the reference tree (REF), nothing copied,
# nothing substituted.  The sources are archived and the driver links against
# the archive, so only the objects its calls reach are linked.  The file list
# is the reference's sequential 3D build (oracle/ref_sources.py reads its
# CMakeLists); files that need libraries this image lacks (UMFPACK, TECIO)
# are left out, and nothing the drivers call reaches them.
# LAPACK (reached through LinAlg.o from THexaAffin, not called on the
# numbering path) is the OpenBLAS that scipy ships in this image, whose
# Fortran symbols carry a scipy_ prefix; the compiler maps the names
# (-Ddgetrf_=scipy_dgetrf_ ...).
#
#   make -f oracle/ref.mk REF=/root/reference -j8
REF ?= /root/reference
OUT := oracle/_ref
CXX := g++
INC := $(addprefix -I$(REF)/include/,General FE Geometry QuadFormulas FE1D FE2D FE3D AMG PBE Refinement System Parallel) \
       -I$(REF)/Examples -I$(REF)/EXT_LIB/tetgen
LAPACK_NAMES := $(foreach f,dgetrf dgetrs dgetri dgemm dsptrd dopgtr dsteqr,-D$(f)_=scipy_$(f)_)
FLAGS := -O2 -w -fpermissive -D__3D__ -D_SEQ $(LAPACK_NAMES)
# the sources of the reference's own sequential 3D build (its CMakeLists)
SRCS := $(shell python3 oracle/ref_sources.py $(REF))
OBJS := $(patsubst $(REF)/src/%.C,$(OUT)/obj/%.o,$(SRCS)) $(OUT)/obj/tetgen.o

all: $(OUT)/ref_numbering $(OUT)/ref_assemble $(OUT)/ref_solve

$(OUT)/obj/%.o: $(REF)/src/%.C
	@mkdir -p $(dir $@)
	$(CXX) $(FLAGS) $(INC) -c $< -o $@

$(OUT)/obj/tetgen.o: $(REF)/EXT_LIB/tetgen/tetgen.cxx
	@mkdir -p $(dir $@)
	$(CXX) -O2 -w -DTETLIBRARY -c $< -o $@

$(OUT)/libparmoon_ref.a: $(OBJS)
	rm -f $@
	ar rcs $@ $(OBJS)

SCIPY_LIBS := $(shell python3 -c "import scipy, os; print(os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), 'scipy.libs'))")
OPENBLAS := $(notdir $(firstword $(wildcard $(SCIPY_LIBS)/libscipy_openblas-*.so)))

$(OUT)/ref_numbering: oracle/ref_numbering.cpp $(OUT)/libparmoon_ref.a
	$(CXX) $(FLAGS) $(INC) oracle/ref_numbering.cpp $(OUT)/libparmoon_ref.a -o $@ \
	  -L$(SCIPY_LIBS) -l:$(OPENBLAS) -Wl,-rpath,$(SCIPY_LIBS)

$(OUT)/ref_assemble: oracle/ref_assemble.cpp oracle/ref_examples.hpp include/pmx_prm.h $(OUT)/libparmoon_ref.a
	$(CXX) $(FLAGS) $(INC) oracle/ref_assemble.cpp $(OUT)/libparmoon_ref.a -o $@ \
	  -L$(SCIPY_LIBS) -l:$(OPENBLAS) -Wl,-rpath,$(SCIPY_LIBS)

$(OUT)/ref_solve: oracle/ref_solve.cpp oracle/ref_examples.hpp include/pmx_prm.h $(OUT)/libparmoon_ref.a
	$(CXX) $(FLAGS) $(INC) oracle/ref_solve.cpp $(OUT)/libparmoon_ref.a -o $@ \
	  -L$(SCIPY_LIBS) -l:$(OPENBLAS) -Wl,-rpath,$(SCIPY_LIBS)
