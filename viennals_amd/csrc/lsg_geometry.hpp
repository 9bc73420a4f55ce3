// Host-side construction of initial bands (defined points only).
#pragma once

#include <cstdint>
#include <vector>

#include "lsg_device.hpp"

namespace lsg {
void sphereBand(int D, const GridInfo &g, const double *origin, double radius, std::vector<int32_t> &idx,
                std::vector<double> &val, int ignoreBcMask = 0);
int planeBand(int D, const GridInfo &g, const double *origin, const double *normal, std::vector<int32_t> &idx,
              std::vector<double> &val);
void boxBand(int D, const GridInfo &g, const double *mn, const double *mx, std::vector<int32_t> &idx,
             std::vector<double> &val, int ignoreBcMask = 0);
}  // namespace lsg
