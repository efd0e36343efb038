# Build of the MI355X (gfx950) semi-naive Datalog engine.
#   make            -- engine (libgdl_hip.so), C++ API + programs (libgdlog.so),
#                      program executables (TC, SG, CSPA, PLEN, DL), generators, CPU oracle
#   make clean
HIPCC    ?= /opt/rocm/bin/hipcc
CC       ?= gcc
PKG      := gdlog-multinode_amd
LIB      := $(PKG)/lib
ARCH     ?= gfx950
HIPFLAGS := --offload-arch=$(ARCH) -O3 -std=c++17 -fPIC -Wall -Iinclude \
            -Xclang -target-feature -Xclang -packed-fp32-ops
CFLAGS   := -O2 -fPIC -Wall -Wextra

ENGINE_SRC := $(wildcard $(PKG)/csrc/*.hip)
ENGINE_OBJ := $(patsubst $(PKG)/csrc/%.hip,$(LIB)/obj/engine_%.o,$(ENGINE_SRC))
API_SRC    := $(wildcard $(PKG)/host/*.cpp) $(wildcard $(PKG)/programs/*.hip) \
              $(PKG)/programs/runner.cpp
API_OBJ    := $(patsubst $(PKG)/%,$(LIB)/obj/%.o,$(API_SRC))
HEADERS    := $(wildcard include/*.h include/gdlog/*.h $(PKG)/csrc/*.hpp $(PKG)/programs/*.h)

all: $(LIB)/libgdl_hip.so $(LIB)/libgdlog.so $(LIB)/libgdl_gen.so oracle/lib/libgdl_oracle.so \
     $(LIB)/TC $(LIB)/SG $(LIB)/CSPA $(LIB)/PLEN $(LIB)/DL

$(LIB)/obj/engine_%.o: $(PKG)/csrc/%.hip $(HEADERS)
	@mkdir -p $(dir $@)
	$(HIPCC) $(HIPFLAGS) -c $< -o $@

$(LIB)/libgdl_hip.so: $(ENGINE_OBJ)
	$(HIPCC) --offload-arch=$(ARCH) -shared -fPIC $^ -o $@

$(LIB)/obj/%.o: $(PKG)/% $(HEADERS)
	@mkdir -p $(dir $@)
	$(HIPCC) $(HIPFLAGS) -x hip -c $< -o $@

$(LIB)/libgdlog.so: $(API_OBJ) $(LIB)/libgdl_hip.so
	$(HIPCC) --offload-arch=$(ARCH) -shared -fPIC $(API_OBJ) -L$(LIB) -lgdl_hip \
	    -Wl,-rpath,'$$ORIGIN' -o $@

$(LIB)/libgdl_gen.so: $(PKG)/gen/gdl_gen.c $(PKG)/gen/gdl_gen.h
	@mkdir -p $(LIB)
	$(CC) $(CFLAGS) -shared $< -o $@ -lm

ORACLE_SRC := oracle/gdl_oracle.c oracle/cpu_tc_omp.c oracle/cpu_souffle.c

oracle/lib/libgdl_oracle.so: $(ORACLE_SRC) oracle/gdl_oracle.h
	@mkdir -p oracle/lib
	$(CC) $(CFLAGS) -fopenmp -shared $(ORACLE_SRC) -o $@

$(LIB)/obj/main_%.o: $(PKG)/programs/main.cpp $(HEADERS)
	@mkdir -p $(dir $@)
	$(HIPCC) $(HIPFLAGS) -DGDL_PROGRAM_$* -x hip -c $< -o $@

$(LIB)/obj/dl_main.o: $(PKG)/programs/dl_main.cpp $(HEADERS)
	@mkdir -p $(dir $@)
	$(HIPCC) $(HIPFLAGS) -x hip -c $< -o $@

$(LIB)/DL: $(LIB)/obj/dl_main.o $(LIB)/libgdlog.so
	$(HIPCC) --offload-arch=$(ARCH) $< -L$(LIB) -lgdlog -lgdl_hip -Wl,-rpath,'$$ORIGIN' -o $@

$(LIB)/%: $(LIB)/obj/main_%.o $(LIB)/libgdlog.so
	$(HIPCC) --offload-arch=$(ARCH) $< -L$(LIB) -lgdlog -lgdl_hip -Wl,-rpath,'$$ORIGIN' -o $@

clean:
	rm -rf $(LIB) oracle/lib

.PHONY: all clean

.SECONDARY:
