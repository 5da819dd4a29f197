# Builds libmcstar_hip.so (gfx950) and the FP64 oracle.
#   make -j8            library + oracle
#   make lib            library only
HIPCC ?= /opt/rocm/bin/hipcc
ARCH ?= gfx950
PKG := mcstar-libs_amd
SRC := $(wildcard $(PKG)/csrc/*.hip)
HDR := $(wildcard $(PKG)/csrc/*.hpp) include/mcstar_hip.h
OBJDIR := $(PKG)/_build
OBJ := $(patsubst $(PKG)/csrc/%.hip,$(OBJDIR)/%.o,$(SRC))
LIB := $(PKG)/lib/libmcstar_hip.so
# packed-fp32 instructions are disabled for all device code
HIPFLAGS := -O3 -std=c++17 -fPIC --offload-arch=$(ARCH) -Xclang -target-feature -Xclang -packed-fp32-ops \
            -Wall -Wno-unused-function -Wno-unused-result -I include

all: lib oracle

lib: $(LIB)

$(OBJDIR)/%.o: $(PKG)/csrc/%.hip $(HDR)
	@mkdir -p $(OBJDIR)
	$(HIPCC) $(HIPFLAGS) -c $< -o $@

$(LIB): $(OBJ)
	@mkdir -p $(dir $@)
	$(HIPCC) -shared -fPIC --offload-arch=$(ARCH) -o $@.tmp $(OBJ)
	mv $@.tmp $@

oracle:
	$(MAKE) -s -C oracle liboracle.so

clean:
	rm -rf $(OBJDIR) $(PKG)/lib

.PHONY: all lib oracle clean
