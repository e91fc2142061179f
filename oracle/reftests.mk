# The reference's own Catch2 tests of the in-scope path, compiled unchanged from where they lie
# under the reference tree against this library's headers (include/) and linked to its libraries,
# into oracle/_ref/ (not committed; the binaries travel to the GPU box with the tree).  They check
# the drop-in boundary: the reference's callers build against include/ and their goldens and
# closed forms pass on the device.  Needs the reference tree (REF) and MPICH (mpi.h, libmpi).
REF ?= /root/reference
ROOT := $(abspath ..)
OUT := _ref
MPI_DIR ?= /opt/conda
TESTS := GF1siteTest GF2siteTest GF4siteTest GF1siteAnomalousTest AndersonTest Anderson2PGFTest Vertex4Test \
         SusceptibilityTest HamiltonianTest 3PSusc1siteTest 3PSusc3siteTest
CXXFLAGS := -std=c++17 -O1 -w -I$(ROOT)/include -I$(REF)/test -idirafter $(MPI_DIR)/include
# libmpi by path, and the system library directory ahead of MPI's in the run path: the MPI
# directory also holds an older libstdc++ that must not replace the system one
SYSLIB := $(dir $(realpath $(shell g++ -print-file-name=libstdc++.so.6)))
LIBS := -L$(ROOT)/pomerol_amd/lib -lpomerol_amd -lpomx $(MPI_DIR)/lib/libmpi.so \
        -Wl,-rpath,$(ROOT)/pomerol_amd/lib:$(SYSLIB):$(MPI_DIR)/lib

all: $(addprefix $(OUT)/,$(TESTS))

$(OUT)/catch2-main.o: $(REF)/test/catch2/catch2-main.cpp
	@mkdir -p $(OUT)
	g++ $(CXXFLAGS) -c $< -o $@

$(OUT)/%: $(REF)/test/%.cpp $(OUT)/catch2-main.o $(wildcard $(ROOT)/include/pomerol/*.hpp) $(ROOT)/pomerol_amd/lib/libpomerol_amd.so
	g++ $(CXXFLAGS) $< $(OUT)/catch2-main.o $(LIBS) -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
