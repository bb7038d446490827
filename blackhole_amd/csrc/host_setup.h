// host_setup.h - host-side scene preparation for the HIP tracer (C++17).
//
// The per-frame scalar work that the reference does on the CPU before its
// thread pool starts (src/render.cpp:399-418) and the one-off host solvers
// its CLI calls (src/main.cpp:399-423 setup_jet, :472-522 setup_neutron_star):
// spectrum table, star catalogue and its cell index, disc constants, camera
// tetrad, jet normalisation, TOV structure.
#pragma once

#include <cstdint>
#include <vector>

#include "params.h"
#include "../../include/bhx.h"

namespace bhx {
namespace host {

// spec:: table of natural-log X, Y, Z over log10 T in [log10 50, log10 3e9]
// (src/spectrum.cpp:97-114): 3 x 4096 doubles, component-major.
const std::vector<double>& bb_table();
void blackbody_xyz(double T, double out[3]);     // src/spectrum.cpp:123-134
void powerlaw_xyz(double alpha, double out[3]);  // src/spectrum.cpp:77-92

// A star catalogue, unit directions + colour temperature + irradiance.
struct Catalogue {
    std::vector<double> x, y, z, T, irr;
};
// The reference's StarField(count, seed) generator (src/scene.h:62-107).
void generate_catalogue(int count, uint64_t seed, Catalogue& cat);
// A synthetic sky of `count` stars: directions uniform on the sphere,
// apparent magnitude m by inverse CDF of N(<m) ~ 10^(0.45 m) on [-1.5, 8],
// B-V ~ N(0.6, 0.4) (Box-Muller) to colour temperature by Ballesteros (2012),
// irradiance F0 10^(-0.4 m) with F0 = 3.64e-9 W m^-2 um^-1 x 0.089 um; a
// splitmix64 stream seeded with `seed`, five draws per star in that order.
void uniform_sky(int count, uint64_t seed, Catalogue& cat);
// Cell index on 360 x 720 lat-lon bins (src/scene.h:173-181): stars sorted by
// cell, ascending index within a cell; rows of 6 doubles x y z irr X/Y Z/Y.
void bin_catalogue(const Catalogue& cat, std::vector<uint32_t>& cell_start,
                   std::vector<double>& rows);

// Scene -> kernel arguments.  Pointers and queue fields are left for the caller.
RaygenArgs raygen_args(const bhx_scene& s);
EventParams event_params(const bhx_scene& s);
TraceArgs trace_args(const bhx_scene& s);
ShadeArgs shade_args(const bhx_scene& s);

// Work-queue order of a W x H frame with S x S samples per pixel: 8 x 8-pixel
// tiles sorted by distance of their centre from the image centre (the photon
// ring's long rays start first), and the first queue slot of each queue tile
// (n_tiles + 1 entries, slot 0 = q_base).  With axis_band, the distance is
// the smaller of that and the distance from the spin axis' image (the line
// through the centre that the camera's roll turns): rays along the axis of a
// conical jet take the most steps inside the jet, so they start first too.
void tile_queue(int W, int H, int S, int32_t q_base, std::vector<int32_t>& order,
                std::vector<int32_t>& start, bool axis_band = false, double roll = 0.0);
// Whether a scene's queue puts the spin axis' image first (the conical jet).
inline bool axis_band(const bhx_scene& s) { return s.jet_enabled && s.jet_model == 1; }

// Kerr helpers used by the set-up code.
double isco_radius(double a, int sense);
double horizon_outer(double a);
double horizon_inner(double a);

}  // namespace host
}  // namespace bhx
