// matv5.hpp -- minimal MATLAB level-5 MAT-file reader/writer (host only).
//
// Covers what the solver exchanges with the reference's tools: numeric arrays of
// any MATLAB numeric class, real or complex, read as double; written as
// uncompressed double (the form rbm.cc writes through libmatio with
// MAT_COMPRESSION_NONE, rbm.cc:406-471, 828-889).  Compressed (miCOMPRESSED)
// elements are read through zlib.  Little-endian files only.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace matv5 {

struct Var {
    std::string name;
    std::vector<int64_t> dims;   // column-major dimensions
    bool complex_ = false;
    int mclass = 0;              // MATLAB class id (6 = double)
    std::vector<double> re, im;  // column-major data
    int64_t numel() const
    {
        int64_t n = 1;
        for (auto d : dims) n *= d;
        return n;
    }
};

// Reads all numeric variables of a file.  Throws std::runtime_error on malformed
// input (with the file name and the reason).
std::vector<Var> read_file(const std::string &path);

// Finds a variable by name (nullptr if absent).
const Var *find(const std::vector<Var> &vars, const std::string &name);

// Writes double variables (real or complex) to a new file.
void write_file(const std::string &path, const std::vector<Var> &vars);

} // namespace matv5
