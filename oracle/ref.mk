# Builds the reference's own svFSI + svFSILS Fortran sources, where they lie
# under /root/reference, with the image's flang, MPICH (/opt/conda) and MKL
# (LAPACK for DGESV in GNNxx), and links oracle/ref_harness.f90 against them.
# Outputs go to oracle/_ref/ only.  TEST INFRASTRUCTURE: this binary writes
# fixtures for tests/golden/ and times the reference on the host; nothing in
# the product path links or runs it.
#
#   make -f oracle/ref.mk -j8          (from the repository root)
#
# The C parts of the reference (SPLIT.c, vtkZpipe.c, remeshTet.cpp,
# trilinos_linear_solver.cpp) are not built: the harness never reaches them
# and vtkZpipe.c needs a CMake-generated header.  MAIN.f is not built either
# (the harness is the main program, and MAIN.f's would pull in the partitioner
# and the VTK writers); the one routine of MAIN.f the built code calls,
# STOPSIM (the error exit of MATFUN's MAT_INV/MAT_EIG), stays unresolved:
# the link ignores that one undefined symbol (check-ref-link lists what is
# undefined), and the harness only feeds non-singular elements.  READFILES.f uses the GNU
# IARGC intrinsic, which flang lacks; it is mapped to the standard
# COMMAND_ARGUMENT_COUNT on the command line, the source is not edited.

REF    ?= /root/reference/Code/Source
FSI    := $(REF)/svFSI
LS     := $(REF)/svFSILS
OUT    := oracle/_ref
OBJ    := $(OUT)/obj
FLANG  ?= /opt/rocm/lib/llvm/bin/flang
MPIDIR ?= /opt/conda
FFLAGS := -O2 -cpp -ffixed-form -ffixed-line-length=132 -I$(FSI) -I$(LS) \
          -I$(MPIDIR)/include -module-dir $(OBJ) -I$(OBJ) \
          "-DIARGC()=COMMAND_ARGUMENT_COUNT()"
F90FLAGS := -O2 -cpp -module-dir $(OBJ) -I$(OBJ)
LIBS   := -L$(MPIDIR)/lib -Wl,-rpath,$(MPIDIR)/lib -lmpifort -lmpi -lmkl_rt

# modules in dependency order (each USEs only modules before it)
MODS := TYPEMOD UTIL MATFUN CHNL LIST COMU ECMOD_DCPLD CEPMOD_AP CEPMOD_BO \
        CEPMOD_FN CEPMOD_TTP CEPMOD MOD ALLFUN
VTK  := vtkTypeParams vtkComMod vtkXMLParser vtkLegacyParser
SKIP := MAIN CONSTS PARAMS_AP PARAMS_BO PARAMS_EC_DCPLD PARAMS_FN PARAMS_TTP $(MODS)
REST := $(filter-out $(SKIP),$(basename $(notdir $(wildcard $(FSI)/*.f))))
LSRC := $(basename $(notdir $(wildcard $(LS)/*.f)))

MOD_O  := $(addprefix $(OBJ)/,$(addsuffix .o,$(MODS)))
VTK_O  := $(addprefix $(OBJ)/,$(addsuffix .o,$(VTK)))
REST_O := $(addprefix $(OBJ)/,$(addsuffix .o,$(REST)))
LS_O   := $(addprefix $(OBJ)/ls_,$(addsuffix .o,$(LSRC)))

all: $(OUT)/svfsi_ref_harness

$(OBJ):
	mkdir -p $(OBJ)

# the module chain is compiled serially, in order
$(OBJ)/modules.stamp: | $(OBJ)
	for m in $(MODS); do $(FLANG) $(FFLAGS) -c $(FSI)/$$m.f -o $(OBJ)/$$m.o || exit 1; done
	for m in $(VTK); do $(FLANG) $(F90FLAGS) -c $(FSI)/$$m.f90 -o $(OBJ)/$$m.o || exit 1; done
	touch $@

$(OBJ)/%.o: $(FSI)/%.f $(OBJ)/modules.stamp
	$(FLANG) $(FFLAGS) -c $< -o $@

$(OBJ)/ls_%.o: $(LS)/%.f | $(OBJ)
	$(FLANG) $(FFLAGS) -c $< -o $@

$(OUT)/libsvfsi_ref.a: $(OBJ)/modules.stamp $(REST_O) $(LS_O)
	rm -f $@
	ar rcs $@ $(MOD_O) $(VTK_O) $(REST_O) $(LS_O)

$(OUT)/svfsi_ref_harness: oracle/ref_harness.f90 $(OUT)/libsvfsi_ref.a
	$(FLANG) -O2 -I$(OBJ) -module-dir $(OBJ) oracle/ref_harness.f90 -o $@ \
	  -Wl,--unresolved-symbols=ignore-all $(OUT)/libsvfsi_ref.a $(LIBS)

# lists the symbols the harness link leaves undefined (expected: stopsim_)
check-ref-link: $(OUT)/libsvfsi_ref.a
	$(FLANG) -O2 -I$(OBJ) -module-dir $(OBJ) oracle/ref_harness.f90 -o $(OUT)/probe.tmp \
	  $(OUT)/libsvfsi_ref.a $(LIBS) 2>&1 | grep -o "undefined symbol: [a-z_0-9]*" | sort -u; \
	rm -f $(OUT)/probe.tmp

clean:
	rm -rf $(OUT)

.PHONY: all clean check-ref-link
