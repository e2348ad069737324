# Builds the reference's own Fortran driver (src/dumses.f90, init, restart,
# output, history, user files) against the device step: the physics files
# (scheme.f90, physics.f90, boundary.f90, fargo.f90, shearingbox.f90 and the
# Riemann solver modules) are left out and dmx_shim.f90 supplies their entry
# points on top of libdmx.so.  Reference sources are compiled where they lie
# (nothing is copied); objects and the binary go to fortran/_build/.
#
#   make -f dumses-hybrid_amd/fortran/dropin.mk [ISO=0|1] [PROBLEM=<dir>]
#
# PROBLEM names a directory of src/problem whose condinit.f90 (and
# history.f90, when present) take the place of the defaults, as the
# reference's VPATH does (src/Makefile.am:43).

REF     ?= /root/reference/src
PKG     := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))/..)
ISO     ?= 0
PROBLEM ?=
FLANG   ?= /opt/rocm/bin/amdflang
OUT     := $(PKG)/fortran/_build/iso$(ISO)$(if $(PROBLEM),_$(PROBLEM),)
DEFS     = -cpp -DMPI=0 -DOACC=0 -DNCCL=0 -DHDF5=0 -DPHDF5=0 -DPNCDF=0 \
           -DCUDA_AWARE_MPI=0 -DNDIM=3 -DGEOM=1 -DCARTESIAN=1 -DCYLINDRICAL=2 \
           -DSPHERICAL=3 -DISO=$(ISO)

PRB      = $(if $(PROBLEM),$(REF)/problem/$(PROBLEM),)
CONDINIT = $(if $(and $(PROBLEM),$(wildcard $(PRB)/condinit.f90)),$(PRB)/condinit.f90,$(REF)/userfiles/condinit.f90)
HISTORY  = $(if $(and $(PROBLEM),$(wildcard $(PRB)/history.f90)),$(PRB)/history.f90,$(REF)/subroutines/history.f90)

PLAIN = $(REF)/modules/commons.f90 $(REF)/modules/file.f90 \
        $(REF)/subroutines/init.f90 $(REF)/subroutines/restart.f90 \
        $(REF)/subroutines/finalize.f90 $(REF)/subroutines/parallel.f90 \
        $(CONDINIT) $(REF)/userfiles/boundana.f90 $(REF)/userfiles/source_term.f90 \
        $(REF)/userfiles/user_init.f90 $(REF)/userfiles/user_restart.f90

$(OUT)/dumses_dmx: $(PLAIN) $(HISTORY) $(REF)/subroutines/output.f90 \
                   $(REF)/userfiles/special.f90 $(REF)/dumses.f90 $(PKG)/fortran/dmx_shim.f90 \
                   $(PKG)/build/libdmx.so
	mkdir -p $(OUT)/obj
	cd $(OUT)/obj && for f in $(PLAIN); do \
	  $(FLANG) -O2 $(DEFS) -module-dir . -c $$f || exit 1; done
	cd $(OUT)/obj && $(FLANG) -O2 $(DEFS) -Dhistory=dmx_ref_history -module-dir . -c $(HISTORY) -o ref_history.o
	cd $(OUT)/obj && $(FLANG) -O2 $(DEFS) -Doutput=dmx_ref_output -module-dir . -c $(REF)/subroutines/output.f90 -o ref_output.o
	cd $(OUT)/obj && $(FLANG) -O2 $(DEFS) -Dspecial=dmx_ref_special -module-dir . -c $(REF)/userfiles/special.f90 -o ref_special.o
	cd $(OUT)/obj && $(FLANG) -O2 $(DEFS) -module-dir . -c $(PKG)/fortran/dmx_shim.f90
	cd $(OUT)/obj && $(FLANG) -O2 $(DEFS) -module-dir . -c $(REF)/dumses.f90
	$(FLANG) -O2 -o $@ $(OUT)/obj/*.o -L$(PKG)/build -ldmx -Wl,-rpath,'$$ORIGIN/../../../build'

clean:
	rm -rf $(PKG)/fortran/_build

.PHONY: clean
