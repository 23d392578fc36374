# Builds the reference Fortran solver (liuspace/Lag2D srccurv-lag) from its
# own sources, read where they lie, into oracle/_ref/ (git-ignored).
# It is a checker only: the parity tests use it to pin the C restatement in
# oracle/ and to write golden fixtures.  Nothing in the product links it.
#
#   make -f oracle/ref.mk            (from the repository root)
#
# Flags: -O2 with -ffp-contract=off so that no fused multiply-add changes the
# reference's rounding; no -fcheck=all (Getrhslagmc.f90:3919 reads
# unkno(:,1,1770) out of bounds on small meshes and would abort).
# The module file constant.mod is generated into _ref/mod; the compiler is run
# from the repository root, so the compiled constant.mod shipped beside the
# reference sources is never read.
REFSRC ?= /root/reference/srccurv-lag
OUT    := oracle/_ref
FC     := amdflang
FFLAGS := -O2 -ffp-contract=off -module-dir $(OUT)/mod -I $(OUT)/mod
UNITS  := dgflo2d Postprocess Preprocess Getrhs Getrhslag Getrhslagmc \
          Reconstruction Limiterlag Getrhslagmctria Getrhslaggauss Lagfvm \
          Getrhslagrz GetBClag Getrhslaggd LinearAlgebra Getsolid GetInitial \
          GetFE Getrhslag_curv Getrhslag_linear WENO
OBJS   := $(addprefix $(OUT)/obj/,$(addsuffix .o,$(UNITS)))

all: $(OUT)/refobjs.a $(OUT)/lag2d_refdump

drivers: $(OUT)/lag2d_drv_cpu $(OUT)/lag2d_drv_gpu $(OUT)/lag2d_drv_gpurhs

$(OUT)/obj/dgflo2d.o: $(REFSRC)/dgflo2d.f90
	@mkdir -p $(OUT)/obj $(OUT)/mod
	$(FC) $(FFLAGS) -c $< -o $@

$(OUT)/obj/%.o: $(REFSRC)/%.f90 $(OUT)/obj/dgflo2d.o
	$(FC) $(FFLAGS) -c $< -o $@

# The reference's program unit (and its C main) are demoted to weak symbols so the dump
# driver (oracle/refdump.f90, written for this project) supplies the program.
$(OUT)/refobjs.a: $(OBJS)
	objcopy --weaken-symbol=_QQmain --weaken-symbol=main $(OUT)/obj/dgflo2d.o $(OUT)/obj/dgflo2d_weak.o
	rm -f $@
	ar rcs $@ $(OUT)/obj/dgflo2d_weak.o $(filter-out $(OUT)/obj/dgflo2d.o,$(OBJS))

$(OUT)/lag2d_refdump: oracle/refdump.f90 $(OUT)/refobjs.a
	$(FC) $(FFLAGS) -o $@ oracle/refdump.f90 $(OUT)/refobjs.a

# The reference's own program twice: as it is (CPU), and with its
# getsolvdg_lag demoted to a weak symbol so that the drop-in of
# lag2d_amd/fortran/getsolvdg_lag_shim.f90 (GPU, through liblag2d.so) takes
# the call.  Used by the integration test of the Fortran boundary.
SHIM := lag2d_amd/fortran
REST := $(filter-out $(OUT)/obj/Getrhslag.o,$(OBJS))

$(OUT)/lag2d_drv_cpu: $(OBJS)
	$(FC) -o $@ $(OBJS)

$(OUT)/obj/lag2d_c.o: $(SHIM)/lag2d_c.f90 $(OUT)/obj/dgflo2d.o
	$(FC) $(FFLAGS) -c $< -o $@

$(OUT)/obj/getsolvdg_lag_shim.o: $(SHIM)/getsolvdg_lag_shim.f90 $(OUT)/obj/lag2d_c.o
	$(FC) $(FFLAGS) -c $< -o $@

$(OUT)/lag2d_drv_gpu: $(OUT)/obj/getsolvdg_lag_shim.o $(OUT)/obj/lag2d_c.o $(OBJS) lag2d_amd/liblag2d.so
	objcopy --weaken-symbol=getsolvdg_lag_ $(OUT)/obj/Getrhslag.o $(OUT)/obj/Getrhslag_weak.o
	$(FC) -o $@ $(OUT)/obj/getsolvdg_lag_shim.o $(OUT)/obj/lag2d_c.o $(OUT)/obj/Getrhslag_weak.o $(REST) \
	    -L lag2d_amd -llag2d -Wl,-rpath,'$$ORIGIN/../../lag2d_amd'

.PHONY: all drivers

# The reference's program with its own getsolvdg_lag (the RK loop on the
# host) and its getrhsdg_lag_mcnew demoted to a weak symbol, so that the
# right-hand-side drop-in of lag2d_amd/fortran/getrhsdg_lag_mcnew_shim.f90
# takes the per-stage call (Getrhslag.f90:1747-1750).
REST_RHS := $(filter-out $(OUT)/obj/Getrhslagmc.o,$(OBJS))

$(OUT)/obj/getrhsdg_lag_mcnew_shim.o: $(SHIM)/getrhsdg_lag_mcnew_shim.f90 $(OUT)/obj/lag2d_c.o
	$(FC) $(FFLAGS) -c $< -o $@

$(OUT)/lag2d_drv_gpurhs: $(OUT)/obj/getrhsdg_lag_mcnew_shim.o $(OUT)/obj/lag2d_c.o $(OBJS) lag2d_amd/liblag2d.so
	objcopy --weaken-symbol=getrhsdg_lag_mcnew_ $(OUT)/obj/Getrhslagmc.o $(OUT)/obj/Getrhslagmc_weak.o
	$(FC) -o $@ $(OUT)/obj/getrhsdg_lag_mcnew_shim.o $(OUT)/obj/lag2d_c.o $(OUT)/obj/Getrhslagmc_weak.o $(REST_RHS) \
	    -L lag2d_amd -llag2d -Wl,-rpath,'$$ORIGIN/../../lag2d_amd'
