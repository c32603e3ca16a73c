// geometry.cpp -- Geometry drop-in (see include/geometry.h).
//
// Builders restate the reference's node sets (src/geometry.cpp:34-175,
// src/geom_object/rectangle.cpp:12-31, src/geom_object/ellipse.cpp:12-33):
// even object sizes shrink by one, the ellipse test is evaluated in unsigned
// 64-bit arithmetic exactly as written there, and the bounds checks throw
// the same std::runtime_error / std::invalid_argument.
#include <algorithm>
#include <fstream>
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

std::vector<size_t> rectangle_nodes(size_t Lx, size_t Ly, size_t xc, size_t yc, size_t Nx) {
    Lx = odd_size(Lx, "x");
    Ly = odd_size(Ly, "y");
    std::vector<size_t> nodes;
    for (size_t ix = xc - Lx / 2; ix <= xc + Lx / 2; ++ix)
        for (size_t iy = yc - Ly / 2; iy <= yc + Ly / 2; ++iy) nodes.push_back(iy * Nx + ix);
    return nodes;
}

std::vector<size_t> ellipse_nodes(size_t Lx, size_t Ly, size_t xc, size_t yc, size_t Nx) {
    Lx = odd_size(Lx, "x");
    Ly = odd_size(Ly, "y");
    const size_t rhs = Lx * Lx * Ly * Ly / 16;
    std::vector<size_t> nodes;
    for (size_t ix = xc - Lx / 2; ix <= xc + Lx / 2; ++ix)
        for (size_t iy = yc - Ly / 2; iy <= yc + Ly / 2; ++iy) {
            const size_t dx = ix - xc, dy = iy - yc;   // wraps mod 2^64; squares are exact mod 2^64
            if (dx * dx * Ly * Ly / 4 + dy * dy * Lx * Lx / 4 <= rhs) nodes.push_back(iy * Nx + ix);
        }
    return nodes;
}

}  // namespace

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
    if (!in) throw std::ios_base::failure("Error opening file: " + fname);
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
    if (!out) throw std::ios_base::failure("Error opening file: " + fname);
    for (size_t y = 0; y < _Ny; ++y) {
        for (size_t x = 0; x < _Nx; ++x) out << geom[y * _Nx + x] << " ";
        out << '\n';
    }
}
