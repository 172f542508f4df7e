// Fixed-size (compile-time) kernels of the 3D operator: dispatch interface.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "kernels.hpp"

namespace b17 {

// Geometry of a supported fixed-size case; the generic plan adopts the same
// split P and pencil width W (and so the same workspace size).
struct StaticGeom {
  int p = 0, w = 0;
  int plane_threads = 0, pencil_threads = 0;
  size_t plane_lds_fwd = 0, plane_lds_inv = 0, pencil_lds = 0;
};

// True if (dtype, N0, N1, N2) has a fixed-size instantiation; fills g.
bool static_supported(int dtype, int n0, int n1, int n2, StaticGeom* g);

// JSON object describing the fixed-size kernels a plan launches (empty if
// the grid has none).
std::string static_describe(int dtype, int n0, int n1, int n2, int pf);

// Launch the three fixed-size kernels (ev: optional 4 timing events).
// pf bit 0 / bit 1: persistent prefetching forward / inverse plane kernel.
// The workspace holds the spectrum rows of NL/2 entries, then the Nyquist
// entries (the same size as the in-row layout of pc).
hipError_t launch_static(int dtype, int n0, int n1, int n2, int mode, const void* in, void* out,
                         void* ws, const PencilParams& pc, const void* twr, const void* twc,
                         int pf, hipStream_t st, hipEvent_t* ev);

}  // namespace b17
