# Builds stock FeCl (the reference decoder) from its sources where they lie under
# /root/reference, plus this repo's golden-vector driver, into oracle/_ref/.
# Nothing is copied out of the reference tree; outputs stay in oracle/_ref/ (git-ignored).
#
#   make -f oracle/ref.mk -j8          (run from the repo root)
#
# The reference's own CMake build is not used: this recipe compiles the LDPC path's
# translation units and the vendored Boost.Serialization sources directly with g++.

REF      ?= /root/reference
OUT      ?= oracle/_ref
CXX      ?= g++
CXXFLAGS ?= -std=c++11 -O3 -DNDEBUG -pthread -fPIC -DBOOST_ALL_NO_LIB -w
INC      := -I$(REF)/libs/include -I$(REF)/src

FECL_SRC := Codec.cpp Trellis.cpp Convolutional.cpp Turbo.cpp Lte3Gpp.cpp Ldpc.cpp DvbS2.cpp \
            detail/Codec.cpp detail/Convolutional.cpp detail/MapDecoder/MapDecoder.cpp \
            detail/MapDecoder/MapDecoderImpl.cpp detail/ViterbiDecoder/ViterbiDecoder.cpp \
            detail/ViterbiDecoder/ViterbiDecoderImpl.cpp detail/Turbo.cpp \
            detail/TurboDecoder/TurboDecoder.cpp detail/TurboDecoder/TurboDecoderImpl.cpp \
            detail/Ldpc.cpp detail/BpDecoder/BpDecoder.cpp detail/BpDecoder/BpDecoderImpl.cpp

SER_SRC  := archive_exception.cpp basic_archive.cpp basic_iarchive.cpp basic_iserializer.cpp \
            basic_oarchive.cpp basic_oserializer.cpp basic_pointer_iserializer.cpp \
            basic_pointer_oserializer.cpp basic_serializer_map.cpp binary_iarchive.cpp \
            binary_oarchive.cpp extended_type_info.cpp extended_type_info_no_rtti.cpp \
            extended_type_info_typeid.cpp void_cast.cpp codecvt_null.cpp stl_port.cpp

FECL_OBJ := $(patsubst %.cpp,$(OUT)/obj/fecl/%.o,$(FECL_SRC))
SER_OBJ  := $(patsubst %.cpp,$(OUT)/obj/ser/%.o,$(SER_SRC))

all: $(OUT)/fecl_golden $(OUT)/fecl_bench

$(OUT)/obj/fecl/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

$(OUT)/obj/ser/%.o: $(REF)/libs/serialization/src/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

$(OUT)/libfecl_ref.a: $(FECL_OBJ) $(SER_OBJ)
	@mkdir -p $(dir $@)
	ar rcs $@ $^

# Golden-vector driver: this repo's own code, linked against stock FeCl.
$(OUT)/fecl_golden: oracle/ref_golden.cpp $(OUT)/libfecl_ref.a
	$(CXX) $(CXXFLAGS) $(INC) -I$(REF)/src/detail/BpDecoder $< $(OUT)/libfecl_ref.a -o $@

# Timing driver: stock FeCl decode on N threads over an LLR file.
$(OUT)/fecl_bench: oracle/ref_bench.cpp $(OUT)/libfecl_ref.a
	$(CXX) $(CXXFLAGS) $(INC) $< $(OUT)/libfecl_ref.a -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
