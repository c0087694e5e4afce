# Builds the reference's own per-step routines (read in place from
# /root/reference, never copied) with ROCm flang into oracle/_ref/, linked with
# the harness oracle/ref_harness.f90 into shared libraries:
#   _ref/libmfa_ref.so      serial build (ziggurat normals), pins the oracle
#   _ref/libmfa_ref_omp.so  OpenMP build (par_ziggurat), CPU-baseline timing
# The reference's default control_simulation.h is used unchanged (BIN_TYPE 2,
# L_BOX 1, WALL 2, SOLVENT 3, THERMOSTAT 0 without DPD_CUT_OFF, BENDING,
# FORCE_SWITCH_ON, POISEUILLE, NO_WARNS).  The flang runtime links statically.
REF    ?= /root/reference
HERE   := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT    := $(HERE)_ref
FLANG  ?= /opt/rocm/llvm/bin/flang
FFLAGS := -cpp -O3 -fPIC -I$(REF)
UNITS  := mfa_common ziggurat par_ziggurat util functions cell_list fluid_fluid dpd_forces intra_molec \
          fluid_wall wall_wall bending constant_force verlet_positions verlet_velocities
SRCS   := $(addprefix $(REF)/,$(addsuffix .f90,$(UNITS))) $(HERE)ref_harness.f90

all: $(OUT)/libmfa_ref.so $(OUT)/libmfa_ref_omp.so

# units are compiled in dependency order (modules first) in one rule per variant
$(OUT)/libmfa_ref.so: $(SRCS)
	mkdir -p $(OUT)/serial
	cd $(OUT)/serial && for u in $(UNITS); do $(FLANG) $(FFLAGS) -c $(REF)/$$u.f90 -o $$u.o || exit 1; done && \
	  $(FLANG) $(FFLAGS) -c $(HERE)ref_harness.f90 -o ref_harness.o && \
	  $(FLANG) -shared -o $@ $(addsuffix .o,$(UNITS)) ref_harness.o

$(OUT)/libmfa_ref_omp.so: $(SRCS)
	mkdir -p $(OUT)/omp
	cd $(OUT)/omp && for u in $(UNITS); do $(FLANG) $(FFLAGS) -fopenmp -c $(REF)/$$u.f90 -o $$u.o || exit 1; done && \
	  $(FLANG) $(FFLAGS) -fopenmp -c $(HERE)ref_harness.f90 -o ref_harness.o && \
	  $(FLANG) -shared -fopenmp -o $@ $(addsuffix .o,$(UNITS)) ref_harness.o

clean:
	rm -rf $(OUT)

.PHONY: all clean
