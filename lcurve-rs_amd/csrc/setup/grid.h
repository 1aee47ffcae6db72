// Surface-element grids (stars, disc, disc edge, bright spot) and their
// brightness. Restated from crates/lcurve/src/{grid,brightness}.rs.
#pragma once

#include "model.h"
#include "roche.h"

namespace lc {

// crates/lcurve/src/grid.rs
std::vector<Point> set_star_grid(const Model& mdl, Star which, bool fine);
std::vector<Point> set_disc_grid(const Model& mdl);
std::vector<Point> set_disc_edge(const Model& mdl, bool outer);
std::vector<Point> set_bright_spot_grid(const Model& mdl);

// crates/lcurve/src/brightness.rs
void set_star_continuum(const Model& mdl, std::vector<Point>& star1, std::vector<Point>& star2);
void set_disc_continuum(double rdisc, double tdisc, double texp, double wave,
                        std::vector<Point>& disc);
void set_edge_continuum(double tedge, double r2, double t2, double absorb, double wave,
                        std::vector<Point>& edge);

// crates/lcurve/src/flux.rs:282-375 (per-curve scalars; not per time)
double comp_gravity1(const Model& mdl, const std::vector<Point>& star1);
double comp_gravity2(const Model& mdl, const std::vector<Point>& star2);
double comp_radius1(const std::vector<Point>& star1);
double comp_radius2(const std::vector<Point>& star2);

}  // namespace lc
