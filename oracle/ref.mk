# Builds the reference CPU tracer (rtow.cxx + its headers) straight from the
# read-only reference tree into oracle/_ref/.  Nothing is copied: g++ reads the
# sources where they lie.  The binary is the oracle's cross-check, never product.
#
#   make -f oracle/ref.mk            (from the repo root)
REF ?= /root/reference
OUT ?= oracle/_ref

all: $(OUT)/rtow

$(OUT)/rtow: $(REF)/rtow.cxx $(REF)/camera.h $(REF)/optics.h $(REF)/ray.h \
             $(REF)/sphere.h $(REF)/thing.h $(REF)/things.h $(REF)/util.h $(REF)/v.h
	mkdir -p $(OUT)
	g++ -std=c++11 -O2 -I$(REF) -o $@ $(REF)/rtow.cxx

.PHONY: all
