/*
 * utils.h -- the small utilities of the reference's include/utils.h
 * (LBMFluids/sc_lbm include/utils.h:20-150, src/utils.cpp:9-14) with the same
 * names and signatures, for drivers and tests written against the reference:
 * subscript <-> linear index conversion, 2-D <-> flat vectors, lower-casing,
 * printing a container, and the relative float comparison the reference's
 * tests use.
 */
#ifndef SCLBM_UTILS_H
#define SCLBM_UTILS_H

#include "common.h"

/// Linear indices y*Nx + x of a list of {x, y} subscripts.
template <typename T>
std::vector<T> sub2ind(std::vector<std::vector<T>>&& subs, const size_t Nx) {
    std::vector<T> lin;
    lin.reserve(subs.size());
    for (const auto& s : subs) lin.push_back(s.at(1) * Nx + s.at(0));
    return lin;
}

template <typename T>
std::vector<T> sub2ind(std::vector<std::vector<T>>& subs, const size_t Nx) {
    std::vector<T> lin;
    lin.reserve(subs.size());
    for (const auto& s : subs) lin.push_back(s.at(1) * Nx + s.at(0));
    return lin;
}

/// Row-major flat vector (Nrow rows of Ncol) to a vector of rows.
template <typename T>
std::vector<std::vector<T>> flat_vector_2_vector_2D(const size_t Ncol, const size_t Nrow,
                                                    const std::vector<T>& flat_vec) {
    std::vector<std::vector<T>> rows(Nrow, std::vector<T>(Ncol));
    for (size_t r = 0; r < Nrow; ++r)
        for (size_t c = 0; c < Ncol; ++c) rows.at(r).at(c) = flat_vec.at(r * Ncol + c);
    return rows;
}

/// Vector of equally long rows to one row-major flat vector.
template <typename T>
std::vector<T> vector_2D_2_flat_vector(const std::vector<std::vector<T>>& vec2D) {
    const size_t ncol = vec2D.at(0).size();
    std::vector<T> flat(vec2D.size() * ncol);
    size_t r = 0;
    for (const auto& row : vec2D) std::copy(row.begin(), row.end(), flat.begin() + (r++) * ncol);
    return flat;
}

/// Lower-case copy of s.
std::string str_to_lower(std::string s);

/// Elements of a container on one line, separated by spaces, after optstr.
template <typename T>
inline void print_elements(const T& coll, const std::string& optstr = "") {
    std::cout << optstr;
    for (const auto& elem : coll) std::cout << elem << ' ';
    std::cout << std::endl;
}

/// |num1 - num2| <= tol * max(1, |num1|, |num2|).
template <typename T>
bool equal_floats(T num1, T num2, T tol) {
    const T scale = std::max({T(1.0), std::fabs(num1), std::fabs(num2)});
    return std::fabs(num1 - num2) <= tol * scale;
}

#endif
