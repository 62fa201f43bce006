# ThreadSanitizer build of the host side: the PTask runtime and the applications whose host-task
# scenarios exercise its threads (runner threads, the dispatch arbiters, per-graph locks, pool
# sweeps).  Host code only, built with g++; the HIP C ABI library is the regular one.
#   make -f tsan.mk          -> ../lib/tsan/libptask.so and ../lib/tsan/bin/{ptk_fdtd,ptk_multigraph,ptk_iteration}
CXX       ?= g++
ROCM      ?= /opt/rocm
ROOT      := $(abspath ../..)
LIBDIR    := $(ROOT)/ptask_amd/lib
OUT       := $(LIBDIR)/tsan
TSANFLAGS := -O1 -g -fPIC -std=c++17 -fsanitize=thread -fno-omit-frame-pointer -D__HIP_PLATFORM_AMD__ \
             -I$(ROCM)/include -I$(ROOT)/include -I$(ROOT)/include/ptask
RT_SRCS   := $(wildcard runtime/*.cpp)
RT_OBJS   := $(patsubst runtime/%.cpp,$(OUT)/obj/%.o,$(RT_SRCS))
TSAN_APPS := ptk_fdtd ptk_multigraph ptk_iteration ptk_cupca

all: $(OUT)/libptask.so $(addprefix $(OUT)/bin/,$(TSAN_APPS))

$(OUT)/obj/%.o: runtime/%.cpp $(wildcard runtime/*.h) $(wildcard $(ROOT)/include/ptask/*.h)
	@mkdir -p $(OUT)/obj
	$(CXX) $(TSANFLAGS) -Iruntime -c -o $@ $<

$(OUT)/libptask.so: $(RT_OBJS) $(LIBDIR)/libptk_hip.so
	$(CXX) -fsanitize=thread -shared -o $@ $(RT_OBJS) -L$(LIBDIR) -lptk_hip -ldl -lpthread \
	    -Wl,-rpath,'$$ORIGIN/..'

$(OUT)/bin/%: apps/%.cpp $(OUT)/libptask.so
	@mkdir -p $(OUT)/bin
	$(CXX) $(TSANFLAGS) -o $@ $< -L$(OUT) -lptask -L$(LIBDIR) -lptk_hip -lpthread \
	    -Wl,-rpath,'$$ORIGIN/..' -Wl,-rpath,'$$ORIGIN/../..'

.PHONY: all
