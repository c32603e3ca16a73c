/*
 * lbm_io.h -- the text reader / writer of the reference
 * (include/io_operations/lbm_io.h:16-118 of LBMFluids/sc_lbm), which drivers
 * and tests use to read back the files Fluid and Geometry write: one row of
 * values per line, separated by the delimiter, 6 significant digits (the
 * stream default).  An unopenable file prints "Error opening file <name>" to
 * stderr and throws std::ios_base::failure, as the reference's FileHandler
 * (src/io_operations/FileHandler.cpp:13-20).
 */
#ifndef SCLBM_LBM_IO_H
#define SCLBM_LBM_IO_H

#include <cerrno>

#include "../common.h"

class LbmIO {
public:
    LbmIO() = default;
    /// name: file; delim: separator written between values; sflag and dims
    /// are kept for the reference's signature (one file, {rows, columns, 0})
    LbmIO(const std::string& name, const std::string& delim, const bool sflag, const std::vector<size_t>& dims)
        : fname(name), delim(delim), single_file(sflag), dimensions(dims) {}

    /// Every line as a row of values (whitespace separated).
    template <typename T>
    std::vector<std::vector<T>> read_vector() const {
        std::ifstream in(fname);
        if (!in.is_open()) fail();
        std::vector<std::vector<T>> rows;
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ss(line);
            std::vector<T> row;
            T v;
            while (ss >> v) row.push_back(v);
            rows.push_back(row);
        }
        return rows;
    }

    /// One line per row, each value followed by the delimiter.
    template <typename T>
    void write_vector(const std::vector<std::vector<T>>& data) const {
        std::ofstream out(fname, std::ios_base::out | std::ios_base::trunc);
        if (!out.is_open()) fail();
        for (const auto& row : data) {
            std::copy(row.begin(), row.end(), std::ostream_iterator<T>(out, delim.c_str()));
            out << '\n';
        }
    }

private:
    std::string fname = "dflt_lbm_io_file.txt";
    std::string delim = " ";
    bool single_file = true;
    std::vector<size_t> dimensions = {0, 0, 0};

    void fail() const {
        std::cerr << "Error opening file " << fname << std::endl;
        throw std::ios_base::failure(std::strerror(errno));
    }
};

#endif
