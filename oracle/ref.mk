# Builds the reference (Fortran, marcjoos-cea/dumses-hybrid) from its
# unmodified sources where they lie, with amdflang, into oracle/_ref/.
# Used only as a checker: to write the golden fixtures under tests/golden/
# and as the timed CPU baseline.  Nothing here is copied into the repository.
#
#   make -f oracle/ref.mk
#
# Targets (all MPI=0, NDIM=3, Cartesian; macros as in src/Makefile.am:30):
#   _ref/iso0/ref_harness      adiabatic, serial     (parity checker)
#   _ref/iso1/ref_harness      isothermal, serial    (parity checker)
#   _ref/iso0_omp/ref_harness  adiabatic, OpenMP -O3 (timed CPU baseline)
# The serial builds are the checker because the reference's OpenMP build
# reads PRIVATE copies of module scratch that the Riemann solvers never
# write (src/modules/solvers/solver.f90:127-128 with :209-242,
# solver_magnetic.f90:114-116 with :208-227), which turns the non-FARGO
# shearing-box terms into NaN; the serial build has the intended values.
# The main program src/dumses.f90 is replaced by oracle/ref_harness.f90,
# which calls the same routines in the same order for a fixed step count.

REF      ?= /root/reference/src
OUT      ?= $(CURDIR)/oracle/_ref
FLANG    ?= /opt/rocm/bin/amdflang
DEFS      = -cpp -DMPI=0 -DOACC=0 -DNCCL=0 -DHDF5=0 -DPHDF5=0 -DPNCDF=0 \
            -DCUDA_AWARE_MPI=0 -DNDIM=3 -DGEOM=1 -DCARTESIAN=1 \
            -DCYLINDRICAL=2 -DSPHERICAL=3

MODSRC = $(REF)/modules/commons.f90 $(REF)/modules/file.f90 \
         $(REF)/modules/solvers/solver.f90 $(REF)/modules/solvers/solver_magnetic.f90
SUBSRC = $(addprefix $(REF)/subroutines/, boundary.f90 fargo.f90 finalize.f90 \
           get_shear.f90 history.f90 init.f90 output.f90 physics.f90 \
           restart.f90 scheme.f90 shearingbox.f90)
USRSRC = $(addprefix $(REF)/userfiles/, boundana.f90 condinit.f90 \
           source_term.f90 special.f90 user_init.f90 user_restart.f90)
HARNESS = $(CURDIR)/oracle/ref_harness.f90
SRCS = $(MODSRC) $(SUBSRC) $(USRSRC) $(HARNESS)

all: $(OUT)/iso0/ref_harness $(OUT)/iso1/ref_harness $(OUT)/iso0_omp/ref_harness

# $(1) = output dir, $(2) = ISO value, $(3) = extra flags
define build
	mkdir -p $(1)/obj
	cd $(1)/obj && for f in $(SRCS); do \
	  $(FLANG) -O2 $(3) $(DEFS) -DISO=$(2) -module-dir . -c $$f || exit 1; done
	$(FLANG) -O2 $(3) -o $(1)/ref_harness $(1)/obj/*.o
endef

$(OUT)/iso0/ref_harness: $(SRCS)
	$(call build,$(OUT)/iso0,0,)
$(OUT)/iso1/ref_harness: $(SRCS)
	$(call build,$(OUT)/iso1,1,)
$(OUT)/iso0_omp/ref_harness: $(SRCS)
	$(call build,$(OUT)/iso0_omp,0,-O3 -fopenmp)

clean:
	rm -rf $(OUT)

.PHONY: all clean
