# Builds oracle/_ref/libfpgm_ref.so from FastPGM's own sources where they lie
# (REF, read-only) plus oracle/ref_driver.cpp.  Runs only where the reference tree
# is present; outputs go to oracle/_ref/ only.  TEST INFRASTRUCTURE ONLY.
#
#   make -f oracle/ref.mk            (from the repository root)
REF ?= /root/reference
OUT = oracle/_ref
CXX ?= g++
CXXFLAGS = -O2 -fPIC -fopenmp -std=c++14 -ffp-contract=off -w -I$(REF)/include
SRCS = $(REF)/src/fastpgm/structure_learning/CellTable.cpp \
       $(REF)/src/fastpgm/ChoiceGenerator.cpp \
       $(REF)/src/fastpgm/Dataset.cpp \
       $(REF)/src/fastpgm/Network.cpp \
       $(REF)/src/fastpgm/Node.cpp \
       $(REF)/src/fastpgm/DiscreteNode.cpp \
       $(REF)/src/fastpgm/ContinuousNode.cpp \
       $(REF)/src/fastpgm/Edge.cpp \
       $(REF)/src/fastpgm/common.cpp \
       $(REF)/src/fastpgm/PotentialTableBase.cpp \
       $(REF)/src/fastpgm/parameter_learning/ParameterLearning.cpp \
       $(REF)/src/fastpgm/structure_learning/BNSLComparison.cpp

all: $(OUT)/libfpgm_ref.so

$(OUT)/libfpgm_ref.so: $(SRCS) oracle/ref_driver.cpp
	mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -shared -o $@ $(SRCS) oracle/ref_driver.cpp -Wl,--no-undefined

clean:
	rm -rf $(OUT)
