// List-mode binning on the GPU (src/binning.cpp, src/planogram.cpp:113-141).
#pragma once
#include <cstdio>
#include <string>
#include "petl_model.h"

namespace petl {

struct BinGeom {
    PlanGeom P;
    float v_m0, v_m1;  // acceptance limits of the planogram
    int nearest;       // 1: nearest u bins (parameters::ALL); 0: clamped bilinear (the reference default)
    int chest_wall;    // 1: list-mode z origin at the chest wall, u1 -= H (the reference default)
};

BinGeom bin_geom(const Planogram& pl, bool nearest, bool chest_wall);

// Adds n coincidences (device arrays blue, pink of n x 3 floats, mm) to the device planogram g in the
// reference's event order; *accepted (device) counts the coincidences the planogram accepts.
void bin_events_device(const BinGeom& B, const float* blue, const float* pink, int64_t n, float* g, BinScratch& W,
                       unsigned long long* accepted, hipStream_t s);

// Reads coincidences from a list-mode file in chunks: ".lst" records of two {float pos[4]; uint64 time}
// events, otherwise the acquisition layout (int64 count; per end point int16 interactions, float pos[3],
// float energy, 18 bytes per further interaction).
class ListModeReader {
public:
    explicit ListModeReader(const char* path);
    ~ListModeReader();
    bool ok() const { return fh_ != nullptr; }
    // fills up to max_events coincidences (x, y, z per end point); returns how many, 0 at the end
    int64_t next(float* blue, float* pink, int64_t max_events);

private:
    std::string path_;
    std::FILE* fh_ = nullptr;
    bool lst_ = false;
    int64_t remaining_ = 0;
};

}  // namespace petl
