// geometry.cpp -- Geometry drop-in (see include/geometry.h).
//
// Builders restate the reference's node sets (src/geometry.cpp:34-175,
// src/geom_object/rectangle.cpp:12-31, src/geom_object/ellipse.cpp:12-33):
// even object sizes shrink by one, the ellipse test is evaluated in unsigned
// 64-bit arithmetic exactly as written there, and the bounds checks throw
// the same std::runtime_error / std::invalid_argument.
#include <algorithm>
#include <cerrno>
#include <cstring>
#include <fstream>
#include <limits>
#include <iostream>
#include <sstream>

#include "geometry.h"
#include "lbm.h"

namespace {

// The reference prints a warning and shrinks an even size by one
// (src/misc_checks.cpp:10-30).
size_t odd_size(size_t L, const char* axis) {
    if (L % 2 == 0) {
        std::cout << "*** WARNING ***\nNumber of nodes in " << axis
                  << " direction is EVEN. Using it will result in an object that is not symmetric."
                  << " Reducing number of nodes by 1." << std::endl;
        return L - 1;
    }
    return L;
}

using Sub = std::vector<size_t>;   // {x, y}

// Node subscripts of a rectangle and of an ellipse centred on (xc, yc), in
// the reference's scan order (x outer, y inner; src/geom_object/rectangle.cpp:12-31,
// src/geom_object/ellipse.cpp:12-33).
std::vector<Sub> rectangle_subs(size_t Lx, size_t Ly, size_t xc, size_t yc) {
    Lx = odd_size(Lx, "x");
    Ly = odd_size(Ly, "y");
    std::vector<Sub> nodes;
    for (size_t ix = xc - Lx / 2; ix <= xc + Lx / 2; ++ix)
        for (size_t iy = yc - Ly / 2; iy <= yc + Ly / 2; ++iy) nodes.push_back({ix, iy});
    return nodes;
}

std::vector<Sub> ellipse_subs(size_t Lx, size_t Ly, size_t xc, size_t yc) {
    Lx = odd_size(Lx, "x");
    Ly = odd_size(Ly, "y");
    const size_t rhs = Lx * Lx * Ly * Ly / 16;
    std::vector<Sub> nodes;
    for (size_t ix = xc - Lx / 2; ix <= xc + Lx / 2; ++ix)
        for (size_t iy = yc - Ly / 2; iy <= yc + Ly / 2; ++iy) {
            const size_t dx = ix - xc, dy = iy - yc;   // wraps mod 2^64; squares are exact mod 2^64
            if (dx * dx * Ly * Ly / 4 + dy * dy * Lx * Lx / 4 <= rhs) nodes.push_back({ix, iy});
        }
    return nodes;
}

std::vector<size_t> linear(const std::vector<Sub>& subs, size_t Nx) {
    std::vector<size_t> lin;
    lin.reserve(subs.size());
    for (const Sub& n : subs) lin.push_back(n[1] * Nx + n[0]);
    return lin;
}

std::vector<size_t> rectangle_nodes(size_t Lx, size_t Ly, size_t xc, size_t yc, size_t Nx) {
    return linear(rectangle_subs(Lx, Ly, xc, yc), Nx);
}

std::vector<size_t> ellipse_nodes(size_t Lx, size_t Ly, size_t xc, size_t yc, size_t Nx) {
    return linear(ellipse_subs(Lx, Ly, xc, yc), Nx);
}

// The node subscripts of a built-in object from the arguments of its add_
// call (src/geometry.cpp:220-240).
std::vector<Sub> built_in_object(const std::string& object_type, const std::vector<size_t>& p) {
    const std::string type = str_to_lower(object_type);
    if (type == "rectangle") return rectangle_subs(p.at(0), p.at(1), p.at(2), p.at(3));
    if (type == "square") return rectangle_subs(p.at(0), p.at(0), p.at(1), p.at(2));
    if (type == "ellipse") return ellipse_subs(p.at(0), p.at(1), p.at(2), p.at(3));
    if (type == "circle") return ellipse_subs(p.at(0), p.at(0), p.at(1), p.at(2));
    throw std::invalid_argument("Non-existing object type. Valid types are: rectangle, square, ellipse, and circle.");
}

// A regular array of copies of one object (src/arrays/regular_array.cpp:11-100).
// The object's extent is its bounding box; the space the array may use along
// an axis is (upper - lower - extent + 1).  Given gaps, the count is
// space / (gap + extent) + 1; given counts, the gap is
// (space + extent) / (count - 1) - extent.  Copy (i, j) is the object shifted
// by i*(extent_x + gap_x) in x and j*(extent_y + gap_y) in y.  All in size_t
// arithmetic, as there.
struct ArrayLayout {
    size_t x0 = std::numeric_limits<size_t>::max(), x1 = 0, y0 = std::numeric_limits<size_t>::max(), y1 = 0;
    size_t len_x = 0, len_y = 0, space_x = 0, space_y = 0;

    ArrayLayout(const std::vector<Sub>& obj, const std::vector<std::vector<size_t>>& bounds) {
        for (const Sub& n : obj) {
            x0 = std::min(x0, n[0]);
            x1 = std::max(x1, n[0]);
            y0 = std::min(y0, n[1]);
            y1 = std::max(y1, n[1]);
        }
        len_x = x1 - x0 + 1;
        len_y = y1 - y0 + 1;
        space_x = bounds.at(0).at(1) - bounds.at(0).at(0) - len_x + 1;
        space_y = bounds.at(1).at(1) - bounds.at(1).at(0) - len_y + 1;
    }

    std::vector<Sub> replicate(const std::vector<Sub>& obj, size_t nx, size_t ny, size_t gx, size_t gy) const {
        const size_t sx = x1 - x0 + gx + 1, sy = y1 - y0 + gy + 1;
        std::vector<Sub> out;
        out.reserve(obj.size() * nx * ny);
        for (const Sub& n : obj)
            for (size_t i = 0; i < nx; ++i)
                for (size_t j = 0; j < ny; ++j) out.push_back({n[0] + sx * i, n[1] + sy * j});
        return out;
    }
};

}  // namespace

std::string str_to_lower(std::string s) {
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

Geometry::Geometry(const Geometry& other) : geom(other.geom), _Nx(other._Nx), _Ny(other._Ny), version(0) {}

Geometry& Geometry::operator=(const Geometry& other) {
    if (this != &other) {
        sclbm_flush();
        geom = other.geom;
        _Nx = other._Nx;
        _Ny = other._Ny;
        ++version;
    }
    return *this;
}

Geometry::~Geometry() {
    sclbm_flush();
    if (dev) scl_geometry_destroy(dev);   // fields built on it keep it alive
}

scl_geometry* Geometry::device() const {
    if (dev && dev_version == version) return dev;
    if (dev) {
        scl_geometry_destroy(dev);
        dev = nullptr;
    }
    if (_Nx == 0 || _Ny == 0) throw std::runtime_error("Geometry is empty");
    scl_geometry* g = nullptr;
    const int rc = scl_geometry_create((int)_Nx, (int)_Ny, geom.data(), &g);
    if (rc != SCL_OK) throw std::runtime_error(std::string("sclbm: ") + scl_last_error());
    dev = g;
    dev_version = version;
    return dev;
}

void Geometry::mark_solid(const std::vector<size_t>& nodes) {
    for (size_t n : nodes) geom.at(n) = 0;
    ++version;
}

void Geometry::add_walls(const size_t dH, const std::string& where) {
    if (where == "y") {
        if (dH > _Nx) throw std::runtime_error("Wall thickness larger than domain width (x)");
        for (size_t iy = 0; iy < _Ny; ++iy)
            for (size_t ix = 0; ix < dH; ++ix) {
                geom.at(iy * _Nx + ix) = 0;
                geom.at(iy * _Nx + (_Nx - 1) - ix) = 0;
            }
    } else if (where == "x") {
        if (dH > _Ny) throw std::runtime_error("Wall thickness larger than domain height (y)");
        for (size_t iy = 0; iy < dH; ++iy)
            for (size_t ix = 0; ix < _Nx; ++ix) {
                geom.at(iy * _Nx + ix) = 0;
                geom.at((_Ny - 1 - iy) * _Nx + ix) = 0;
            }
    } else {
        throw std::invalid_argument("No options for position argument: " + where);
    }
    ++version;
}

void Geometry::add_rectangle(const size_t Lx, const size_t Ly, const size_t xc, const size_t yc) {
    check_object_bounds(Lx, Ly, xc, yc, "Rectangle");
    const std::vector<size_t> obj = rectangle_nodes(Lx, Ly, xc, yc, _Nx);
    check_constructed_object(obj, "rectangle");
    mark_solid(obj);
}

void Geometry::add_square(const size_t Ls, const size_t xc, const size_t yc) {
    check_object_bounds(Ls, Ls, xc, yc, "Square");
    const std::vector<size_t> obj = rectangle_nodes(Ls, Ls, xc, yc, _Nx);
    check_constructed_object(obj, "square");
    mark_solid(obj);
}

void Geometry::add_ellipse(const size_t Lx, const size_t Ly, const size_t xc, const size_t yc) {
    check_object_bounds(Lx, Ly, xc, yc, "Ellipse");
    const std::vector<size_t> obj = ellipse_nodes(Lx, Ly, xc, yc, _Nx);
    check_constructed_object(obj, "ellipse");
    mark_solid(obj);
}

void Geometry::add_circle(const size_t D, const size_t xc, const size_t yc) {
    check_object_bounds(D, D, xc, yc, "Circle");
    const std::vector<size_t> obj = ellipse_nodes(D, D, xc, yc, _Nx);
    check_constructed_object(obj, "circle");
    mark_solid(obj);
}

void Geometry::add_array(const std::vector<size_t>& object_properties,
                         const std::vector<std::vector<size_t>>& array_bounds, const size_t dx, const size_t dy,
                         const std::string& object_type, const size_t) {
    const std::vector<Sub> obj = built_in_object(object_type, object_properties);
    const ArrayLayout lay(obj, array_bounds);
    const size_t nx = lay.space_x / (dx + lay.len_x) + 1, ny = lay.space_y / (dy + lay.len_y) + 1;
    ++version;
    for (size_t n : linear(lay.replicate(obj, nx, ny, dx, dy), _Nx)) geom.at(n) = 0;
}

void Geometry::add_array(const std::vector<size_t>& object_properties,
                         const std::vector<std::vector<size_t>>& array_bounds,
                         const std::vector<size_t>& object_numbers, const std::string& object_type, const size_t) {
    const std::vector<Sub> obj = built_in_object(object_type, object_properties);
    const ArrayLayout lay(obj, array_bounds);
    const size_t nx = object_numbers.at(0), ny = object_numbers.at(1);
    // the reference divides by (count - 1) unchecked; one object per axis has no gap
    if (nx < 2 || ny < 2) throw std::invalid_argument("add_array: at least two objects in x and in y");
    const size_t gx = (lay.space_x + lay.len_x) / (nx - 1) - lay.len_x;
    const size_t gy = (lay.space_y + lay.len_y) / (ny - 1) - lay.len_y;
    ++version;
    for (size_t n : linear(lay.replicate(obj, nx, ny, gx, gy), _Nx)) geom.at(n) = 0;
}

void Geometry::check_object_bounds(const size_t Lx, const size_t Ly, const size_t xc, const size_t yc,
                                   const std::string& name) const {
    if (((int)xc - (int)Lx / 2) < 0 || ((int)yc - (int)Ly / 2) < 0)
        throw std::runtime_error(name + " lower bounds do not fit the domain.");
    if ((xc + Lx / 2) >= _Nx || (yc + Ly / 2) >= _Ny)
        throw std::runtime_error(name + " upper bounds do not fit the domain.");
}

void Geometry::check_constructed_object(const std::vector<size_t>& obj, const std::string& name) const {
    if (!obj.empty() && *std::max_element(obj.begin(), obj.end()) >= _Nx * _Ny)
        throw std::runtime_error("Constructed " + name + " upper bounds do not fit the domain.");
}

void Geometry::read(const std::string& fname) {
    std::ifstream in(fname);
    if (!in) {
        std::cerr << "Error opening file " << fname << std::endl;
        throw std::ios_base::failure(std::strerror(errno));
    }
    std::vector<std::vector<int>> rows;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::vector<int> r;
        int v;
        while (ss >> v) r.push_back(v);
        rows.push_back(r);
    }
    _Nx = rows.at(0).size();
    _Ny = rows.size();
    geom.assign(_Nx * _Ny, 0);
    for (size_t y = 0; y < _Ny; ++y)
        for (size_t x = 0; x < _Nx; ++x) geom[y * _Nx + x] = rows[y].at(x);
    ++version;
}

void Geometry::write(const std::string& fname) const {
    std::ofstream out(fname, std::ios_base::out | std::ios_base::trunc);
    if (!out) {
        std::cerr << "Error opening file " << fname << std::endl;
        throw std::ios_base::failure(std::strerror(errno));
    }
    for (size_t y = 0; y < _Ny; ++y) {
        for (size_t x = 0; x < _Nx; ++x) out << geom[y * _Nx + x] << " ";
        out << '\n';
    }
}
