# Builds libmcstar_hip.so (gfx950) and the FP64 oracle.
#   make -j8            library + oracle
#   make lib            library only
HIPCC ?= /opt/rocm/bin/hipcc
ARCH ?= gfx950
PKG := mcstar-libs_amd
SRC := $(wildcard $(PKG)/csrc/*.hip)
OBJDIR := $(PKG)/_build
OBJ := $(patsubst $(PKG)/csrc/%.hip,$(OBJDIR)/%.o,$(SRC))
LIB := $(PKG)/lib/libmcstar_hip.so
# packed-fp32 instructions are disabled for all device code; multiply-adds are fused only where
# the source writes fma (mfma), so every kernel built from the same expressions rounds alike
HIPFLAGS := -O3 -std=c++17 -fPIC --offload-arch=$(ARCH) -Xclang -target-feature -Xclang -packed-fp32-ops -ffp-contract=off \
            -Wall -Wno-unused-function -Wno-unused-result -I include -MMD -MP

all: lib oracle driver sim

# the simulator driver (mcstar-libs_amd/sim), one build per precision
SIM := $(PKG)/bin/mcs_sim_f32 $(PKG)/bin/mcs_sim_f64
sim: $(SIM)

$(PKG)/bin/mcs_sim_f32: $(PKG)/sim/mcs_sim.cpp $(LIB) $(wildcard include/*/*.hpp include/*/*/*.hpp)
	@mkdir -p $(dir $@)
	$(HIPCC) -O2 -std=c++17 -DFLOAT=float -I include $< -o $@ -L$(PKG)/lib -lmcstar_hip -lz -Wl,-rpath,'$$ORIGIN/../lib'

$(PKG)/bin/mcs_sim_f64: $(PKG)/sim/mcs_sim.cpp $(LIB) $(wildcard include/*/*.hpp include/*/*/*.hpp)
	@mkdir -p $(dir $@)
	$(HIPCC) -O2 -std=c++17 -DFLOAT=double -I include $< -o $@ -L$(PKG)/lib -lmcstar_hip -lz -Wl,-rpath,'$$ORIGIN/../lib'

# C++ drop-in driver of the kept host API (tests/cpp), one build per precision
DRV := tests/cpp/bin/kept_api_driver_f32 tests/cpp/bin/kept_api_driver_f64 \
       tests/cpp/bin/kept_api_exp_driver_f32 tests/cpp/bin/kept_api_exp_driver_f64
driver: $(DRV)

tests/cpp/bin/kept_api_exp_driver_%: tests/cpp/kept_api_exp_driver.cpp $(LIB) $(wildcard include/*/*.hpp include/*/*/*.hpp)
	@mkdir -p $(dir $@)
	$(HIPCC) -O2 -std=c++17 -DFLOAT=$(if $(filter f64,$*),double,float) -I include $< -o $@ -L$(PKG)/lib -lmcstar_hip -lz -Wl,-rpath,'$$ORIGIN/../../../$(PKG)/lib'

tests/cpp/bin/kept_api_driver_f32: tests/cpp/kept_api_driver.cpp $(LIB) $(wildcard include/*/*.hpp include/*/*/*.hpp)
	@mkdir -p $(dir $@)
	$(HIPCC) -O2 -std=c++17 -DFLOAT=float -I include $< -o $@ -L$(PKG)/lib -lmcstar_hip -lz -Wl,-rpath,'$$ORIGIN/../../../$(PKG)/lib'

tests/cpp/bin/kept_api_driver_f64: tests/cpp/kept_api_driver.cpp $(LIB) $(wildcard include/*/*.hpp include/*/*/*.hpp)
	@mkdir -p $(dir $@)
	$(HIPCC) -O2 -std=c++17 -DFLOAT=double -I include $< -o $@ -L$(PKG)/lib -lmcstar_hip -lz -Wl,-rpath,'$$ORIGIN/../../../$(PKG)/lib'

lib: $(LIB)

# each object depends on the headers it includes (dependency files written by the compiler)
$(OBJDIR)/%.o: $(PKG)/csrc/%.hip
	@mkdir -p $(OBJDIR)
	$(HIPCC) $(HIPFLAGS) -c $< -o $@

-include $(OBJ:.o=.d)

$(LIB): $(OBJ) $(PKG)/csrc/exports.map
	@mkdir -p $(dir $@)
	$(HIPCC) -shared -fPIC --offload-arch=$(ARCH) -Wl,--version-script=$(PKG)/csrc/exports.map -o $@.tmp $(OBJ)
	mv $@.tmp $@

oracle:
	$(MAKE) -s -C oracle liboracle.so

clean:
	rm -rf $(OBJDIR) $(PKG)/lib $(PKG)/bin tests/cpp/bin

.PHONY: all lib oracle driver clean
