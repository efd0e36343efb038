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

# ---- AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code
# (CPU only): the oracle and the CPU baselines under gcc, and the C++ API's host
# code (Datalog front end, LIE, relations, RA operators) under hipcc with the
# sanitizers on the host side only.  `make sanitize` builds and runs both.
SAN      := oracle/lib/asan
SANFLAGS := -fsanitize=address,undefined -fno-sanitize-recover=all -fno-omit-frame-pointer -g -O1
HOSTSAN  := -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined \
            -Xarch_host -fno-sanitize-recover=all -Xarch_host -fno-omit-frame-pointer
SAN_HOST := $(patsubst %,$(SAN)/host_%.o,datalog lie relation ra_ops)

$(SAN)/oracle_asan: oracle/sanitize/oracle_asan_main.c $(ORACLE_SRC) oracle/gdl_oracle.h
	@mkdir -p $(SAN)
	$(CC) $(SANFLAGS) -fopenmp oracle/sanitize/oracle_asan_main.c $(ORACLE_SRC) -o $@ -lm

$(SAN)/host_%.o: $(PKG)/host/%.cpp $(HEADERS)
	@mkdir -p $(SAN)
	$(HIPCC) --offload-arch=$(ARCH) -O1 -g -std=c++17 -fPIC -Iinclude \
	    -Xclang -target-feature -Xclang -packed-fp32-ops $(HOSTSAN) -x hip -c $< -o $@

$(SAN)/frontend_asan: oracle/sanitize/frontend_asan_main.cpp $(SAN_HOST) $(LIB)/libgdl_hip.so
	$(HIPCC) -O1 -g -std=c++17 -Iinclude $(HOSTSAN) -c oracle/sanitize/frontend_asan_main.cpp \
	    -o $(SAN)/frontend_main.o
	$(HIPCC) --offload-arch=$(ARCH) $(HOSTSAN) $(SAN)/frontend_main.o $(SAN_HOST) -L$(LIB) \
	    -lgdl_hip -Wl,-rpath,'$$ORIGIN/../../../$(LIB)' -o $@

sanitize: $(SAN)/oracle_asan $(SAN)/frontend_asan
	ASAN_OPTIONS=detect_leaks=1 $(SAN)/oracle_asan
	ASAN_OPTIONS=detect_leaks=0 $(SAN)/frontend_asan $(PKG)/programs/dl/*.dl

.PHONY: sanitize
