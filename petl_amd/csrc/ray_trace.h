// Analytic ray tracing through geometric solids: line integrals of a geometric phantom (ray_trace) and the
// stopping power of the scintillator modules (stopping_power), one value per planogram bin.
//
// Reference: src/phantom.cpp:502-640, 1120-1141 (solids), src/analytic_ray_tracing_gpu.cu:48-997 (tracer),
// src/petl.cpp:81-152 (drivers).
#pragma once
#include <vector>
#include "petl_common.h"

namespace petl {

// One solid, laid out as the reference's geometricObject / geometricSolid.
struct Solid {
    int type;  // 0 ellipsoid, 1 parallelepiped, 2/3/4 cylinder x/y/z, 5/6/7 cone x/y/z
    float c[3], r[3];
    float val;
    float A[9];          // rotation (row-major), applied as A (x - c)
    float clip[6][4];    // clipping planes (normal, offset): keep n.x >= offset
    int rotated;
    int nclip;
    float cone[2];       // axial clip range of a cone about its apex
    // bounding sphere of the unclipped solid, for culling (radius < 0: always test)
    float bc[3], br;
};

// geometricObject::init: false for a null centre / radius or an unknown type.  A and clip may be null.
bool solid_init(int type, const float* c, const float* r, float val, const float* A, const float* clip, Solid& out);
// phantom::scale_phantom (src/phantom.cpp:1120-1141)
void solids_scale(std::vector<Solid>& s, float sx, float sy, float sz);

// Every bin of one planogram (device array g, overwritten): stopping power (1 - e^-a)(1 - e^-b) of the
// two half rays, or the line integral, averaged over oversampling^2 sub-rays (oversampling clamped to an
// odd number in [1, 11]).  d_solids: n solids in device memory.
void ray_trace_device(const PlanGeom& P, const Solid* d_solids, int n, int oversampling, bool stopping, float* g,
                      hipStream_t s);

}  // namespace petl
