# Builds oracle/_ref/ref_numbering (test infrastructure): the reference's own
# mesh, refinement, FE-space and matrix-structure code compiled with g++ from
# the sources where they lie under the reference tree (REF), nothing copied,
# nothing substituted.  The sources are archived and the driver links against
# the archive, so only the objects its calls reach are linked.  Files the
# sequential 3D build cannot compile here without libraries this image lacks
# (UMFPACK, TECIO) are left out; nothing the driver calls reaches them.
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
SRCS := $(wildcard $(REF)/src/General/*.C $(REF)/src/Geometry/*.C $(REF)/src/Refinement/*.C \
          $(REF)/src/QuadFormulas/*.C $(REF)/src/FE/*.C)
SKIP := DirectSolver Output3D Output2D 2d3d_FDM Assemble3D_bkup RKV_FDM TCD2D TimeConvDiff2D Windtunnel_3d4d
SRCS := $(filter-out $(foreach s,$(SKIP),%/$(s).C),$(SRCS))
OBJS := $(patsubst $(REF)/src/%.C,$(OUT)/obj/%.o,$(SRCS)) $(OUT)/obj/tetgen.o

all: $(OUT)/ref_numbering

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
