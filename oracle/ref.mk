# Build of the reference's own pixel-level routines (Fortran 77) into
# oracle/_ref/, compiled from the sources where they lie under the
# reference tree.  TEST INFRASTRUCTURE ONLY: used to pin the C restatement
# (oracle/ref32.c) and as the CPU baseline; nothing in the product loads it.
#
#     make -f oracle/ref.mk            (from the repository root)
#
# The library leaves the reference's message routines (ERRREPOR, ERRTRACE,
# MSGPUT: they need the reference's string layer, which it does not ship)
# and FFTPACK (CFFTI/CFFTF/CFFTB/RFFTI/RFFTF/RFFTB: only reached for sizes
# that are not powers of two) undefined.  It is loaded with lazy binding,
# so those are never resolved unless called; the callers only use inputs
# that reach neither.
REF ?= /root/reference/source
FLANG ?= /opt/rocm/lib/llvm/bin/flang
OUT ?= oracle/_ref
FFLAGS ?= -O2
# -fno-automatic: the routines rely on static locals (FFTX keeps its bit-
# reversal and trig tables, not SAVEd, between calls of one length), as on
# the platforms the reference was written for
FPP = -cpp -ffixed-form -fno-automatic -DCOMP_GNU -DMACH_LINUX -DOS_LINUX -I$(REF)/include

SRC = $(REF)/grd/grdh2d.f $(REF)/grd/grddeh2d.f $(REF)/grd/grdqh2d.f $(REF)/grd/grdcf.f \
      $(REF)/grd/grdh23d.f $(REF)/grd/grddeh23.f $(REF)/grd/grdqh23d.f \
      $(REF)/grd/grdsf.f $(REF)/grd/grdsze.f $(REF)/grd/grdccf.f $(REF)/grd/grduwt2d.f \
      $(REF)/grd/grdruw2d.f $(REF)/pix/pixcor2d.f $(REF)/pix/pixcorx2.f $(REF)/fft/fftr2d.f \
      $(REF)/fft/fftr1d2.f $(REF)/fft/fftr1d.f $(REF)/fft/fftr.f $(REF)/fft/fftx1d.f \
      $(REF)/fft/fftx.f $(REF)/fft/fftpwr2.f $(REF)/fft/fftrnp2.f $(REF)/fft/fftxnp2.f \
      $(REF)/vis/visselpi.f

all: $(OUT)/libsderef.so $(OUT)/libsderef_O0.so

$(OUT)/libsderef.so: $(SRC)
	mkdir -p $(OUT)
	$(FLANG) $(FPP) $(FFLAGS) -fPIC -shared -o $@ $(SRC)

# the reference's own build flags compile without optimisation
$(OUT)/libsderef_O0.so: $(SRC)
	mkdir -p $(OUT)
	$(FLANG) $(FPP) -O0 -fPIC -shared -o $@ $(SRC)

clean:
	rm -f $(OUT)/libsderef.so $(OUT)/libsderef_O0.so

.PHONY: all clean
