/*
 * geometry.h -- drop-in for the reference's Geometry (include/geometry.h:37-286
 * of LBMFluids/sc_lbm): the solid mask of an Nx x Ny lattice, 1 = fluid,
 * 0 = solid, row-major, with the same builders, bounds checks, accessors and
 * text I/O.  The mask is mirrored to the GPU (one bit per site) the first
 * time a device operation needs it and again after every change.
 *
 */
#ifndef SCLBM_GEOMETRY_H
#define SCLBM_GEOMETRY_H

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.h"
#include "sc_lbm_amd.h"
#include "utils.h"

class Geometry {
public:
    /// Empty geometry, to be filled by read() (include/geometry.h:49)
    Geometry() = default;
    /// Nx x Ny lattice filled with fluid (include/geometry.h:58-59)
    Geometry(const size_t Nx, const size_t Ny) : _Nx(Nx), _Ny(Ny) { geom.resize(_Nx * _Ny, 1); }
    Geometry(const Geometry& other);
    Geometry& operator=(const Geometry& other);
    ~Geometry();

    /// Solid walls of thickness dH: "x" = rows at the bottom and top,
    /// "y" = columns at the left and right (src/geometry.cpp:34-69)
    void add_walls(const size_t dH = 1, const std::string& where = "y");

    /// Objects (src/geometry.cpp:76-147): even sizes are reduced by one
    void add_rectangle(const size_t Lx, const size_t Ly, const size_t xc, const size_t yc);
    void add_square(const size_t Ls, const size_t xc, const size_t yc);
    void add_ellipse(const size_t Lx, const size_t Ly, const size_t xc, const size_t yc);
    void add_circle(const size_t D, const size_t xc, const size_t yc);

    /// Regular arrays of a built-in object (src/geometry.cpp:182-240,
    /// src/arrays/regular_array.cpp:11-100).  object_properties are the
    /// arguments of the matching add_ call ({Lx, Ly, xc, yc} for "rectangle"
    /// and "ellipse", {L, xc, yc} for "square" and "circle", case-insensitive);
    /// array_bounds = {{x lower, x upper}, {y lower, y upper}}.  The first form
    /// takes the gaps between neighbouring objects, the second the numbers of
    /// objects in x and y.  alpha (staggering) is unused, as in the reference.
    void add_array(const std::vector<size_t>& object_properties, const std::vector<std::vector<size_t>>& array_bounds,
                   const size_t dx, const size_t dy, const std::string& object_type, const size_t alpha = 0);
    void add_array(const std::vector<size_t>& object_properties, const std::vector<std::vector<size_t>>& array_bounds,
                   const std::vector<size_t>& object_numbers, const std::string& object_type,
                   const size_t alpha = 0);

    /// Bounds checks (src/geometry.cpp:150-175); throw std::runtime_error
    void check_object_bounds(const size_t Lx, const size_t Ly, const size_t xc, const size_t yc,
                             const std::string& name) const;
    void check_constructed_object(const std::vector<size_t>& obj, const std::string& name) const;

    void set_node_fluid(const size_t xi, const size_t yi) { geom.at(yi * _Nx + xi) = 1; ++version; }
    void set_node_solid(const size_t xi, const size_t yi) { geom.at(yi * _Nx + xi) = 0; ++version; }

    /// Text I/O: one row per y, values separated by spaces (src/geometry.cpp:247-282)
    void read(const std::string& fname);
    void write(const std::string& fname) const;

    const int operator()(const size_t xi, const size_t yi) const { return geom.at(yi * _Nx + xi); }
    const int operator()(const size_t ind) const { return geom.at(ind); }

    size_t Nx() const { return _Nx; }
    size_t Ny() const { return _Ny; }
    const std::vector<int>& get_geom() const { return geom; }

    /// Device mirror of the current mask (created or refreshed on demand).
    scl_geometry* device() const;

private:
    std::vector<int> geom;
    size_t _Nx = 0, _Ny = 0;
    unsigned long version = 0;
    mutable scl_geometry* dev = nullptr;
    mutable unsigned long dev_version = ~0ul;

    void mark_solid(const std::vector<size_t>& nodes);
};

#endif
