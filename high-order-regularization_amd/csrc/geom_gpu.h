// GPU geometry drivers (geom_gpu.hip).  Host pointers in; outputs to the host
// (surface, nearest data) or to device buffers (stencils).
#pragma once

#include "hor.h"

namespace hor {

// Generate_Surface (surface.cpp:13-71): count in *n; filled when cap >= *n.
int surface_generate_gpu(hor_ctx* ctx, const hor_body& B, double h, double pou_deg, int cap, int* n,
                         double* x, double* nrml, double* area, int* plane);

// Compute_nearest_data (surface.cpp:450-526) of nt targets against ns
// quadrature points with Monge planes splane.
int nearest_data_gpu(hor_ctx* ctx, const hor_body& B, double h, double DEL, int ns, const double* sx,
                     const int* splane, int nt, const double* tx, int* nrst, double* x0, double* n0, double* b,
                     int* i123, double* f0);

// 6x6 Monge stencils of nt targets (nearest data nrst, x0, i123) on a surface
// of ns points sx with centre ctr: st_dev[36 nt] source indices, [k1][k2];
// co_dev[12 nt] the 6-point Lagrange weights in the two in-plane directions.
int stencil_build(hor_ctx* ctx, const double ctr[3], double h, int ns, const double* sx, int nt, const int* nrst,
                  const double* x0, const int* i123, int* st_dev, double* co_dev);

}  // namespace hor
