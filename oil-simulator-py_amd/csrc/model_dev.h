// Device-side grid helpers shared by the IMPES step kernels (impes_model.hip) and
// their reverse sweep (adjoint.hip).
#pragma once

#include "impes_model.h"

namespace impes {
namespace model_dev {

constexpr int kBlock = 256;
constexpr double kDay = 86400.0;

struct GridDev {
  int nx, ny, nz;
  double dx, dy, dz;
  long long sx, sy;
  int gravity;
  int env_nx;  // > 0: independent reservoirs of env_nx x-layers each, stacked along x
};

#define CELL_LOOP(c, n) for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < (n); c += (long long)gridDim.x * blockDim.x)

__device__ __forceinline__ void ijk(const GridDev& G, long long c, int& i, int& j, int& k) {
  k = (int)(c % G.nz);
  const long long t = c / G.nz;
  j = (int)(t % G.ny);
  i = (int)(t / G.ny);
}

inline GridDev grid_dev(const Grid& g, bool gravity, int env_nx) {
  GridDev G;
  G.nx = g.nx; G.ny = g.ny; G.nz = g.nz;
  G.dx = g.dx; G.dy = g.dy; G.dz = g.dz;
  G.sx = (long long)g.ny * g.nz;
  G.sy = g.nz;
  G.gravity = gravity ? 1 : 0;
  G.env_nx = env_nx;
  return G;
}

__device__ __forceinline__ double mx12(double a) { return fmax(a, 1e-12); }

}  // namespace model_dev
}  // namespace impes
